/*
 * or_core.c -- oracle for the RK4 step (TEST INFRASTRUCTURE ONLY; see
 * oracle.h).  Every routine names the reference lines it restates.  Loops
 * whose iterations are independent may run in any order (and in parallel);
 * reductions and particle deposits keep the reference's order so the
 * rounding matches it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define OR_PI 3.141592653589793 /* params.h:5 */
#define NT1 17
#define NT0 19

struct or_state {
  or_params p;
  int nx, ny, nz, bs, NX, NY, NZ;
  long NXNY, n;
  double dx, dy, dz, idx, idy, idz, idx2, idy2, idz2;
  double R, P0par, Tp, mp, Vcell, xh, yh, zh;
  double A, gx, gy, gz, I0;
  int Iteration;
  double pre[4], post[4];
  double *t1[NT1][3];
  double *t0[NT0];
  double *cc, *rhs_new, *pd[3], *fft;
  int num_iteration;
  long total_iterations;  /* Poisson iterations summed over every solve */
  double P0, P0_int, P0_new, P0_np1, dP0_dt, met, T_cur;
  int num_timestep, rk;
  int np, cap, has_tau, has_eps;
  double *q[7][4];
  double *aux[10];
  int *flag;
  double *scratch[4];
  int *bk_plane, *bk_idx, *bk_start, *bk_cnt;  /* deposit buckets by z-plane */
  int bk_cap, bk_threads;
};

#define G(s, i, j, k) ((long)(k) * (s)->NXNY + (long)(j) * (s)->NX + (long)(i))

static void load_int(or_state *s, double *f, const double *in);
static void store_int(const or_state *s, const double *f, double *out);

/* ------------------------------------------------------------------ */
/* allocation                                                          */
/* ------------------------------------------------------------------ */
static double *fld(const or_state *s) { return (double *)calloc((size_t)s->n, sizeof(double)); }

or_state *or_create(const or_params *p, int np_capacity)
{
  or_state *s = (or_state *)calloc(1, sizeof(or_state));
  s->p = *p;
  s->nx = p->nx; s->ny = p->ny; s->nz = p->nz; s->bs = p->bs;
  s->NX = p->nx + 2 * p->bs; s->NY = p->ny + 2 * p->bs; s->NZ = p->nz + 2 * p->bs;
  s->NXNY = (long)s->NX * s->NY;
  s->n = s->NXNY * s->NZ;
  /* gridsize.cpp:23-25, tensor0.cpp:12-32 */
  s->dx = p->lx / p->nx; s->dy = p->ly / p->ny; s->dz = p->lz / p->nz;
  s->idx = 1. / s->dx; s->idy = 1. / s->dy; s->idz = 1. / s->dz;
  s->idx2 = 1. / (s->dx * s->dx); s->idy2 = 1. / (s->dy * s->dy); s->idz2 = 1. / (s->dz * s->dz);
  s->Vcell = s->dx * s->dy * s->dz;                 /* gridsize.h:31 */
  s->xh = (double)p->nx * s->dx; s->yh = (double)p->ny * s->dy; s->zh = (double)p->nz * s->dz; /* gridsize.cpp:67-77 */
  /* params.cpp:61-72 */
  s->R = p->cp - p->cv;
  s->P0par = p->rho0 * p->t0 * s->R;
  s->Tp = p->rhop * p->dp * p->dp / (18 * p->mu0);
  s->mp = p->rhop * OR_PI * p->dp * p->dp * p->dp / 6.;
  s->Iteration = p->iteration1; s->A = p->a1;
  s->gx = p->gx1; s->gy = p->gy1; s->gz = p->gz1; s->I0 = p->i01;
  /* grid.cpp:18-25 */
  s->pre[0] = 0.5; s->pre[1] = 0.5; s->pre[2] = 1.; s->pre[3] = 1. / 6.;
  s->post[0] = 1. / 6.; s->post[1] = 1. / 3.; s->post[2] = 1. / 3.; s->post[3] = 1. / 6.;
  for (int a = 0; a < NT1; ++a) for (int c = 0; c < 3; ++c) s->t1[a][c] = fld(s);
  for (int a = 0; a < NT0; ++a) s->t0[a] = fld(s);
  s->cc = fld(s); s->rhs_new = fld(s);
  for (int c = 0; c < 3; ++c) s->pd[c] = fld(s);
  for (int c = 0; c < 4; ++c) s->scratch[c] = fld(s);
  s->fft = (double *)calloc(2 * (size_t)p->nx * p->ny * p->nz, sizeof(double));
  s->cap = np_capacity > 0 ? np_capacity : 1;
  for (int a = 0; a < 7; ++a) for (int b = 0; b < 4; ++b) s->q[a][b] = (double *)calloc((size_t)s->cap, sizeof(double));
  for (int a = 0; a < 10; ++a) s->aux[a] = (double *)calloc((size_t)s->cap, sizeof(double));
  s->flag = (int *)calloc((size_t)s->cap, sizeof(int));
  return s;
}

void or_destroy(or_state *s)
{
  if (!s) return;
  for (int a = 0; a < NT1; ++a) for (int c = 0; c < 3; ++c) free(s->t1[a][c]);
  for (int a = 0; a < NT0; ++a) free(s->t0[a]);
  free(s->cc); free(s->rhs_new);
  for (int c = 0; c < 3; ++c) free(s->pd[c]);
  for (int c = 0; c < 4; ++c) free(s->scratch[c]);
  free(s->fft);
  for (int a = 0; a < 7; ++a) for (int b = 0; b < 4; ++b) free(s->q[a][b]);
  for (int a = 0; a < 10; ++a) free(s->aux[a]);
  free(s->flag);
  free(s->bk_plane); free(s->bk_idx); free(s->bk_start); free(s->bk_cnt);
  free(s);
}

int or_nthreads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

static double **T1(or_state *s, int id) { return s->t1[id]; }
static double *T0(or_state *s, int id) { return s->t0[id - OR_P]; }

/* ------------------------------------------------------------------ */
/* ghost exchange semantics on one rank                                */
/* ------------------------------------------------------------------ */
/* communicator.cpp:108-216: x sweep (interior y,z), then y (all x, interior
 * z), then z (all x,y).  Each ghost becomes its periodic image. */
static void fill(const or_state *s, double *f)
{
  int bs = s->bs, NX = s->NX, NY = s->NY, NZ = s->NZ;
  int nx = s->nx, ny = s->ny, nz = s->nz;
#pragma omp parallel for schedule(static)
  for (int k = bs; k < NZ - bs; ++k)
    for (int j = bs; j < NY - bs; ++j)
      for (int g = 0; g < bs; ++g) {
        f[G(s, NX - bs + g, j, k)] = f[G(s, bs + g, j, k)];
        f[G(s, g, j, k)] = f[G(s, g + nx, j, k)];
      }
#pragma omp parallel for schedule(static)
  for (int k = bs; k < NZ - bs; ++k)
    for (int g = 0; g < bs; ++g)
      for (int i = 0; i < NX; ++i) {
        f[G(s, i, NY - bs + g, k)] = f[G(s, i, bs + g, k)];
        f[G(s, i, g, k)] = f[G(s, i, g + ny, k)];
      }
#pragma omp parallel for schedule(static)
  for (int g = 0; g < bs; ++g)
    for (int j = 0; j < NY; ++j)
      for (int i = 0; i < NX; ++i) {
        f[G(s, i, j, NZ - bs + g)] = f[G(s, i, j, bs + g)];
        f[G(s, i, j, g)] = f[G(s, i, j, g + nz)];
      }
}

/* communicator.cpp:228-317: sweeps z (all x,y), y (all x, interior z), x
 * (interior y,z); each ghost deposit is ADDED to its periodic owner. */
static void fold(const or_state *s, double *f)
{
  int bs = s->bs, NX = s->NX, NY = s->NY, NZ = s->NZ;
  for (int i = 0; i < NX; ++i)
    for (int j = 0; j < NY; ++j)
      for (int g = 0; g < bs; ++g) {
        double lo = f[G(s, i, j, g)], hi = f[G(s, i, j, NZ - bs + g)];
        f[G(s, i, j, NZ - 2 * bs + g)] += lo;
        f[G(s, i, j, bs + g)] += hi;
      }
  for (int i = 0; i < NX; ++i)
    for (int g = 0; g < bs; ++g)
      for (int k = bs; k < NZ - bs; ++k) {
        double lo = f[G(s, i, g, k)], hi = f[G(s, i, NY - bs + g, k)];
        f[G(s, i, NY - 2 * bs + g, k)] += lo;
        f[G(s, i, bs + g, k)] += hi;
      }
  for (int g = 0; g < bs; ++g)
    for (int j = bs; j < NY - bs; ++j)
      for (int k = bs; k < NZ - bs; ++k) {
        double lo = f[G(s, g, j, k)], hi = f[G(s, NX - bs + g, j, k)];
        f[G(s, NX - 2 * bs + g, j, k)] += lo;
        f[G(s, bs + g, j, k)] += hi;
      }
}

/* ------------------------------------------------------------------ */
/* element-wise operations over all elements incl. ghosts              */
/* (tensor0.cpp:54-201)                                                */
/* ------------------------------------------------------------------ */
#define EW_LOOP for (long e = 0; e < s->n; ++e)
static void ew_set(const or_state *s, double *o, double a) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] = a; }
static void ew_copy(const or_state *s, double *o, const double *a) { memcpy(o, a, sizeof(double) * (size_t)s->n); }
static void ew_add_s(const or_state *s, double *o, double a) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] += a; }
static void ew_sub_s(const or_state *s, double *o, double a) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] -= a; }
static void ew_mul(const or_state *s, double *o, const double *a) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] *= a[e]; }
static void ew_sub(const or_state *s, double *o, const double *a) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] -= a[e]; }
static void ew_add(const or_state *s, double *o, const double *a) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] += a[e]; }
static void ew_mul_s(const or_state *s, double *o, double a) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] *= a; }
/* PlusEqual_Mult(double a, tensor0 b): o += a*b */
static void ew_pem(const or_state *s, double *o, double a, const double *b) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] += a * b[e]; }
/* Equal_Divide(tensor0 a, tensor0 b): o = a/b */
static void ew_div(const or_state *s, double *o, const double *a, const double *b) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] = a[e] / b[e]; }
/* Equal_Divide(double a, tensor0 b): o = a/b */
static void ew_sdiv(const or_state *s, double *o, double a, const double *b) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] = a / b[e]; }
/* Equal_Mult(tensor0 a, double b) */
static void ew_mulc(const or_state *s, double *o, const double *a, double b) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] = a[e] * b; }
/* Equal_LinComb(a,T,b,R): o = a*T + b*R */
static void ew_lc2(const or_state *s, double *o, double a, const double *T, double b, const double *Rr) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] = a * T[e] + b * Rr[e]; }
/* Equal_LinComb(a,T,b,R,c,S): o = a*T + b*R + c*S */
static void ew_lc3(const or_state *s, double *o, double a, const double *T, double b, const double *Rr, double c, const double *S) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] = a * T[e] + b * Rr[e] + c * S[e]; }

/* ------------------------------------------------------------------ */
/* stencils on the interior, followed by a ghost fill                  */
/* (tensor0.cpp:204-296,354-362; tensor1.cpp:259-311)                   */
/* ------------------------------------------------------------------ */
#define IN_LOOP(s) \
  _Pragma("omp parallel for schedule(static)") \
  for (int k = (s)->bs; k < (s)->NZ - (s)->bs; ++k) \
    for (int j = (s)->bs; j < (s)->NY - (s)->bs; ++j) \
      for (int i = (s)->bs; i < (s)->NX - (s)->bs; ++i)

/* S5: Equal_Div_F2C (tensor0.cpp:249-256) */
static void div_f2c(or_state *s, double *o, double *const *T)
{
  long sx = 1, sy = s->NX, sz = s->NXNY;
  IN_LOOP(s) {
    long c = G(s, i, j, k);
    o[c] = (T[0][c + sx] - T[0][c]) * s->idx + (T[1][c + sy] - T[1][c]) * s->idy +
           (T[2][c + sz] - T[2][c]) * s->idz;
  }
  fill(s, o);
}

/* S6: Equal_Del2 (tensor0.cpp:258-266), d2dx2 (:234-247) */
static void del2(or_state *s, double *o, const double *f)
{
  long sx = 1, sy = s->NX, sz = s->NXNY;
  IN_LOOP(s) {
    long c = G(s, i, j, k);
    o[c] = (f[c + sx] - 2 * f[c] + f[c - sx]) * s->idx2 +
           (f[c + sy] - 2 * f[c] + f[c - sy]) * s->idy2 +
           (f[c + sz] - 2 * f[c] + f[c - sz]) * s->idz2;
  }
  fill(s, o);
}

/* S7: tensor1::Equal_Grad_C2F (tensor1.cpp:259-270) */
static void grad_c2f(or_state *s, double *const *o, const double *f)
{
  long sx = 1, sy = s->NX, sz = s->NXNY;
  IN_LOOP(s) {
    long c = G(s, i, j, k);
    o[0][c] = (f[c] - f[c - sx]) * s->idx;
    o[1][c] = (f[c] - f[c - sy]) * s->idy;
    o[2][c] = (f[c] - f[c - sz]) * s->idz;
  }
  fill(s, o[0]); fill(s, o[1]); fill(s, o[2]);
}

/* S4: I?_C2F (tensor0.cpp:288-296) into o, then fill */
static void i_c2f_axis(or_state *s, double *o, const double *f, int axis)
{
  long st = axis == 0 ? 1 : (axis == 1 ? s->NX : s->NXNY);
  IN_LOOP(s) {
    long c = G(s, i, j, k);
    o[c] = (f[c] + f[c - st]) * 0.5;
  }
  fill(s, o);
}
/* S4: I?_F2C (tensor0.cpp:268-286) */
static void i_f2c_axis(or_state *s, double *o, const double *f, int axis)
{
  long st = axis == 0 ? 1 : (axis == 1 ? s->NX : s->NXNY);
  IN_LOOP(s) {
    long c = G(s, i, j, k);
    o[c] = (f[c] + f[c + st]) * 0.5;
  }
  fill(s, o);
}
/* S8: tensor1::Equal_I_C2F(tensor0) (tensor1.cpp:278-283) */
static void i_c2f(or_state *s, double *const *o, const double *f)
{
  i_c2f_axis(s, o[0], f, 0); i_c2f_axis(s, o[1], f, 1); i_c2f_axis(s, o[2], f, 2);
}
/* S8: tensor1::Equal_I{x,y,z}_C2F(tensor1) (tensor1.cpp:292-311) */
static void iaxis_c2f_t1(or_state *s, double *const *o, double *const *T, int axis)
{
  i_c2f_axis(s, o[0], T[0], axis); i_c2f_axis(s, o[1], T[1], axis); i_c2f_axis(s, o[2], T[2], axis);
}
/* S9: tensor0::Equal_I_F2C(tensor1) (tensor0.cpp:354-362) */
static void i_f2c_mean(or_state *s, double *o, double *const *T)
{
  double third = 1. / 3.;
  long sx = 1, sy = s->NX, sz = s->NXNY;
  IN_LOOP(s) {
    long c = G(s, i, j, k);
    o[c] = third * ((T[0][c] + T[0][c + sx]) * 0.5 + (T[1][c] + T[1][c + sy]) * 0.5 +
                    (T[2][c] + T[2][c + sz]) * 0.5);
  }
  fill(s, o);
}

/* ------------------------------------------------------------------ */
/* reductions (tensor0.cpp:364-396, 466-478): interior, i->j->k order  */
/* ------------------------------------------------------------------ */
static double red_sum(const or_state *s, const double *f)
{
  double acc = 0;
  for (int i = s->bs; i < s->NX - s->bs; ++i)
    for (int j = s->bs; j < s->NY - s->bs; ++j)
      for (int k = s->bs; k < s->NZ - s->bs; ++k) acc += f[G(s, i, j, k)];
  return acc;
}
static double red_mean(const or_state *s, const double *f)
{
  return red_sum(s, f) / (double)(s->nx * s->ny * s->nz); /* sum()/size_tot() */
}
static double red_max_abs(const or_state *s, const double *f)
{
  double m = fabs(f[G(s, s->bs, s->bs, s->bs)]);
  for (int k = s->bs; k < s->NZ - s->bs; ++k)
    for (int j = s->bs; j < s->NY - s->bs; ++j)
      for (int i = s->bs; i < s->NX - s->bs; ++i) {
        double v = fabs(f[G(s, i, j, k)]);
        if (v > m) m = v;
      }
  return m;
}
/* tensor0::make_mean_zero (tensor0.cpp:380-384) */
static void make_mean_zero(or_state *s, double *f) { double m = red_mean(s, f); ew_sub_s(s, f, m); }

/* ------------------------------------------------------------------ */
/* field access                                                        */
/* ------------------------------------------------------------------ */
static double *field_ptr(or_state *s, int id, int comp)
{
  if (id >= 0 && id < NT1 && comp >= 0 && comp < 3) return s->t1[id][comp];
  if (id >= OR_P && id < OR_P + NT0) return s->t0[id - OR_P];
  return NULL;
}

int or_set_field(or_state *s, int id, int comp, const double *in)
{
  double *f = field_ptr(s, id, comp);
  if (!f) return -1;
  long m = 0;
  for (int k = s->bs; k < s->NZ - s->bs; ++k)
    for (int j = s->bs; j < s->NY - s->bs; ++j)
      for (int i = s->bs; i < s->NX - s->bs; ++i) f[G(s, i, j, k)] = in[m++];
  fill(s, f);
  return 0;
}

int or_get_field(const or_state *s, int id, int comp, double *out)
{
  double *f = field_ptr((or_state *)s, id, comp);
  if (!f) return -1;
  long m = 0;
  for (int k = s->bs; k < s->NZ - s->bs; ++k)
    for (int j = s->bs; j < s->NY - s->bs; ++j)
      for (int i = s->bs; i < s->NX - s->bs; ++i) out[m++] = f[G(s, i, j, k)];
  return 0;
}

int or_get_field_ghosted(const or_state *s, int id, int comp, double *out)
{
  double *f = field_ptr((or_state *)s, id, comp);
  if (!f) return -1;
  memcpy(out, f, sizeof(double) * (size_t)s->n);
  return 0;
}

/* ------------------------------------------------------------------ */
/* particles                                                           */
/* ------------------------------------------------------------------ */
enum { QX = 0, QY, QZ, QU, QV, QW, QT };

int or_set_particles(or_state *s, int n, const double *x, const double *y, const double *z,
                     const double *u, const double *v, const double *w, const double *T,
                     const int *flag, const double *tau, const double *eps)
{
  if (n > s->cap || n < 0) return -1;
  const double *src[7] = {x, y, z, u, v, w, T};
  for (int a = 0; a < 7; ++a) memcpy(s->q[a][OR_BASE], src[a], sizeof(double) * (size_t)n);
  memcpy(s->flag, flag, sizeof(int) * (size_t)n);
  s->has_tau = tau != NULL; s->has_eps = eps != NULL;
  if (tau) memcpy(s->aux[OR_AUX_TAU], tau, sizeof(double) * (size_t)n);
  if (eps) memcpy(s->aux[OR_AUX_EPS], eps, sizeof(double) * (size_t)n);
  s->np = n;
  return 0;
}

int or_get_np(const or_state *s) { return s->np; }

int or_get_particles(const or_state *s, int stage, double *const *q, int *flag)
{
  if (stage < 0 || stage > 3) return -1;
  for (int a = 0; a < 7; ++a)
    if (q && q[a]) memcpy(q[a], s->q[a][stage], sizeof(double) * (size_t)s->np);
  if (flag) memcpy(flag, s->flag, sizeof(int) * (size_t)s->np);
  return 0;
}

int or_get_particle_aux(const or_state *s, int which, double *out)
{
  if (which < 0 || which >= 10) return -1;
  memcpy(out, s->aux[which], sizeof(double) * (size_t)s->np);
  return 0;
}

int or_set_particle_aux(or_state *s, int which, const double *in)
{
  if (which < 0 || which >= 10) return -1;
  memcpy(s->aux[which], in, sizeof(double) * (size_t)s->np);
  return 0;
}

/* particle.cpp:106-134 (single rank: il = jl = kl = 0) */
static int x2i(const or_state *s, double x) { return (int)(x * s->idx) + s->bs; }
static int y2j(const or_state *s, double y) { return (int)(y * s->idy) + s->bs; }
static int z2k(const or_state *s, double z) { return (int)(z * s->idz) + s->bs; }

/* P1: particle::gas2part_velocity (particle.cpp:340-360) at x_int */
void or_p_gather_velocity(or_state *s)
{
  double **U = T1(s, OR_U);
  double *xi = s->q[QX][OR_INT], *yi = s->q[QY][OR_INT], *zi = s->q[QZ][OR_INT];
#pragma omp parallel for schedule(static)
  for (int p = 0; p < s->np; ++p) {
    int ii = x2i(s, xi[p]), jj = y2j(s, yi[p]), kk = z2k(s, zi[p]);
    double xl = xi[p] * s->idx - (int)(xi[p] * s->idx);
    double yl = yi[p] * s->idy - (int)(yi[p] * s->idy);
    double zl = zi[p] * s->idz - (int)(zi[p] * s->idz);
    s->aux[OR_AUX_UG][p] = xl * U[0][G(s, ii + 1, jj, kk)] + (1 - xl) * U[0][G(s, ii, jj, kk)];
    s->aux[OR_AUX_VG][p] = yl * U[1][G(s, ii, jj + 1, kk)] + (1 - yl) * U[1][G(s, ii, jj, kk)];
    s->aux[OR_AUX_WG][p] = zl * U[2][G(s, ii, jj, kk + 1)] + (1 - zl) * U[2][G(s, ii, jj, kk)];
  }
}

/* P2: particle::update_position (particle.cpp:314-338) */
static void update_position(or_state *s)
{
  double dt = s->p.dt, post = s->post[s->rk], pre = s->pre[s->rk];
  for (int a = QX; a <= QZ; ++a) {
    double *np1 = s->q[a][OR_NP1], *nw = s->q[a][OR_NEW], *b = s->q[a][OR_BASE];
    const double *ui = s->q[a + 3][OR_INT];
#pragma omp parallel for schedule(static)
    for (int p = 0; p < s->np; ++p) np1[p] += dt * post * ui[p];
    if (s->rk != 3) {
#pragma omp parallel for schedule(static)
      for (int p = 0; p < s->np; ++p) nw[p] = b[p] + dt * pre * ui[p];
    } else memcpy(nw, np1, sizeof(double) * (size_t)s->np);
  }
}

/* P3: particle::update_velocity (particle.cpp:362-399); per-particle tau is
 * the library's generalisation of the scalar Tp. */
static void update_velocity(or_state *s)
{
  double iTp = s->Tp != 0 ? 1. / s->Tp : 0.;
  double dt = s->p.dt, post = s->post[s->rk], pre = s->pre[s->rk];
  double g[3] = {(double)s->p.particle_gravity * s->gx, (double)s->p.particle_gravity * s->gy,
                 (double)s->p.particle_gravity * s->gz};
  double *gas[3] = {s->aux[OR_AUX_UG], s->aux[OR_AUX_VG], s->aux[OR_AUX_WG]};
  double *rhs[3] = {s->aux[OR_AUX_RHS_U], s->aux[OR_AUX_RHS_V], s->aux[OR_AUX_RHS_W]};
  for (int a = 0; a < 3; ++a) {
    const double *ui = s->q[QU + a][OR_INT];
#pragma omp parallel for schedule(static)
    for (int p = 0; p < s->np; ++p) {
      double itp = s->has_tau ? 1. / s->aux[OR_AUX_TAU][p] : iTp;
      rhs[a][p] = (gas[a][p] - ui[p]) * itp;
    }
  }
  for (int a = 0; a < 3; ++a) {
    double *np1 = s->q[QU + a][OR_NP1], *nw = s->q[QU + a][OR_NEW], *b = s->q[QU + a][OR_BASE];
#pragma omp parallel for schedule(static)
    for (int p = 0; p < s->np; ++p) np1[p] += dt * post * (rhs[a][p] + g[a]);
    if (s->rk != 3) {
#pragma omp parallel for schedule(static)
      for (int p = 0; p < s->np; ++p) nw[p] = b[p] + dt * pre * (rhs[a][p] + g[a]);
    } else memcpy(nw, np1, sizeof(double) * (size_t)s->np);
  }
}

/* P4: particle::update_Temp (particle.cpp:477-492); per-particle epsilon is
 * the library's generalisation of the scalar epsilonp. */
static void update_temp(or_state *s)
{
  double dt = s->p.dt, post = s->post[s->rk], pre = s->pre[s->rk];
  double iS = 1. / (s->mp * s->p.cvp);
  double *np1 = s->q[QT][OR_NP1], *nw = s->q[QT][OR_NEW], *b = s->q[QT][OR_BASE];
  const double *rt = s->aux[OR_AUX_RHS_T];
#pragma omp parallel for schedule(static)
  for (int p = 0; p < s->np; ++p) {
    double epsp = s->has_eps ? s->aux[OR_AUX_EPS][p] : s->p.epsilonp;
    double alpha = 0.25 * epsp * OR_PI * s->p.dp * s->p.dp * s->I0;
    np1[p] += dt * post * iS * (rt[p] + alpha);
  }
  if (s->rk != 3) {
#pragma omp parallel for schedule(static)
    for (int p = 0; p < s->np; ++p) {
      double epsp = s->has_eps ? s->aux[OR_AUX_EPS][p] : s->p.epsilonp;
      double alpha = 0.25 * epsp * OR_PI * s->p.dp * s->p.dp * s->I0;
      nw[p] = b[p] + dt * pre * iS * (rt[p] + alpha);
    }
  } else memcpy(nw, np1, sizeof(double) * (size_t)s->np);
}

/* CIC index/weight helper (particle.cpp:409-420, 437-448): ii=round(x/dx),
 * xl = x/dx - ii + 0.5, then made local with +bs. */
static void cic(const or_state *s, double x, double y, double z, int *ii, int *jj, int *kk,
                double *xl, double *yl, double *zl)
{
  int i = (int)(round(x * s->idx)), j = (int)(round(y * s->idy)), k = (int)(round(z * s->idz));
  *xl = x * s->idx - i + 0.5; *yl = y * s->idy - j + 0.5; *zl = z * s->idz - k + 0.5;
  *ii = i + s->bs; *jj = j + s->bs; *kk = k + s->bs;
}

/* Deposits run in parallel without changing the order in which any cell
 * receives its contributions.  The particles are bucketed by the z-plane of
 * their stencil's base cell (a stable counting sort: each thread counts and
 * places one contiguous index range), and each thread then owns whole target
 * planes: for target plane t it merges, in increasing particle index, the
 * buckets of the two base planes whose stencils reach t and adds only the
 * contributions that land in t.  Every cell therefore sums its particles in
 * the order of the reference's serial loop (particle.cpp:494-654), so the
 * result is the same bit for bit at any thread count; on P ranks the
 * reference itself splits the loop the same way, one subdomain per rank. */
static int nthreads_now(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* plane[p] in [-1, NZ] is stored shifted by one; buckets live in bk_idx
 * between bk_start[b] and bk_start[b + 1], b = plane + 1 */
static void bucket_planes(or_state *s)
{
  int np = s->np, nb = s->NZ + 2, T = nthreads_now();
  if (T > np / 4096 + 1) T = np / 4096 + 1;
  if (T < 1) T = 1;
  if (s->bk_threads < T) {
    free(s->bk_cnt);
    s->bk_cnt = (int *)calloc((size_t)T * (nb + 1), sizeof(int));
    s->bk_threads = T;
  }
  if (!s->bk_start) s->bk_start = (int *)calloc((size_t)nb + 1, sizeof(int));
  int *cnt = s->bk_cnt, *pl = s->bk_plane, *idx = s->bk_idx, *st = s->bk_start;
#pragma omp parallel num_threads(T)
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    long lo = (long)np * t / T, hi = (long)np * (t + 1) / T;
    int *c = cnt + (size_t)t * (nb + 1);
    for (int b = 0; b <= nb; ++b) c[b] = 0;
    for (long p = lo; p < hi; ++p) c[pl[p] + 1]++;
#pragma omp barrier
#pragma omp single
    {
      int run = 0;
      for (int b = 0; b < nb; ++b) {
        st[b] = run;
        for (int u = 0; u < T; ++u) {
          int k = cnt[(size_t)u * (nb + 1) + b];
          cnt[(size_t)u * (nb + 1) + b] = run;
          run += k;
        }
      }
      st[nb] = run;
    }
    for (long p = lo; p < hi; ++p) idx[c[pl[p] + 1]++] = (int)p;
  }
}

static void bucket_alloc(or_state *s)
{
  if (s->bk_cap < s->np) {
    free(s->bk_plane); free(s->bk_idx);
    s->bk_cap = s->cap > s->np ? s->cap : s->np;
    s->bk_plane = (int *)malloc(sizeof(int) * (size_t)s->bk_cap);
    s->bk_idx = (int *)malloc(sizeof(int) * (size_t)s->bk_cap);
  }
}

/* next particle of the merge of buckets b0 and b1 (increasing index), -1 at the end */
#define MERGE_NEXT(ia, ea, ib, eb, idx) \
  ((ia) < (ea) && ((ib) >= (eb) || (idx)[ia] < (idx)[ib]) ? (idx)[(ia)++] : ((ib) < (eb) ? (idx)[(ib)++] : -1))

/* the contributions of particle p's cloud-in-cell stencil (base cell
 * (ii, jj, kk), planes kk and kk - 1) that land in plane t (t < 0: all), in
 * the reference's expression order (particle.cpp:522-555, 620-654) */
static inline void cic_one(or_state *s, double *A, const double *x, const double *y, const double *z,
                           const double *w, int p, int t)
{
  int ii, jj, kk; double xl, yl, zl;
  cic(s, x[p], y[p], z[p], &ii, &jj, &kk, &xl, &yl, &zl);
  int lo = t < 0 || kk == t, hi = t < 0 || kk - 1 == t;
  if (w) {
    double r = w[p];
    if (lo) {
      A[G(s, ii, jj, kk)] += xl * yl * zl * r;
      A[G(s, ii - 1, jj, kk)] += (1 - xl) * yl * zl * r;
      A[G(s, ii, jj - 1, kk)] += xl * (1 - yl) * zl * r;
      A[G(s, ii - 1, jj - 1, kk)] += (1 - xl) * (1 - yl) * zl * r;
    }
    if (hi) {
      A[G(s, ii, jj, kk - 1)] += xl * yl * (1 - zl) * r;
      A[G(s, ii - 1, jj, kk - 1)] += (1 - xl) * yl * (1 - zl) * r;
      A[G(s, ii, jj - 1, kk - 1)] += xl * (1 - yl) * (1 - zl) * r;
      A[G(s, ii - 1, jj - 1, kk - 1)] += (1 - xl) * (1 - yl) * (1 - zl) * r;
    }
  } else {
    if (lo) {
      A[G(s, ii, jj, kk)] += (xl * yl * zl);
      A[G(s, ii - 1, jj, kk)] += (1 - xl) * yl * zl;
      A[G(s, ii, jj - 1, kk)] += xl * (1 - yl) * zl;
      A[G(s, ii - 1, jj - 1, kk)] += (1 - xl) * (1 - yl) * zl;
    }
    if (hi) {
      A[G(s, ii, jj, kk - 1)] += xl * yl * (1 - zl);
      A[G(s, ii - 1, jj, kk - 1)] += (1 - xl) * yl * (1 - zl);
      A[G(s, ii, jj - 1, kk - 1)] += xl * (1 - yl) * (1 - zl);
      A[G(s, ii - 1, jj - 1, kk - 1)] += (1 - xl) * (1 - yl) * (1 - zl);
    }
  }
}

/* cloud-in-cell deposit of w (NULL: weight 1) at positions x,y,z into A */
static void deposit_cic(or_state *s, double *A, const double *x, const double *y, const double *z,
                        const double *w)
{
  if (nthreads_now() == 1) {
    for (int p = 0; p < s->np; ++p) cic_one(s, A, x, y, z, w, p, -1);
    return;
  }
  bucket_alloc(s);
#pragma omp parallel for schedule(static)
  for (int p = 0; p < s->np; ++p) s->bk_plane[p] = (int)(round(z[p] * s->idz)) + s->bs;
  bucket_planes(s);
  const int *idx = s->bk_idx, *st = s->bk_start;
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < s->NZ; ++t) {
    int ia = st[t + 1], ea = st[t + 2], ib = st[t + 2], eb = st[t + 3];  /* base t, base t + 1 */
    for (int p; (p = MERGE_NEXT(ia, ea, ib, eb, idx)) >= 0;) cic_one(s, A, x, y, z, w, p, t);
  }
}

/* P5: particle::part2gas_concentration (particle.cpp:522-555) */
void or_p_deposit_concentration(or_state *s)
{
  double *A = T0(s, OR_C);
  ew_set(s, A, 0.);
  deposit_cic(s, A, s->q[QX][OR_INT], s->q[QY][OR_INT], s->q[QZ][OR_INT], NULL);
  fold(s, A);
}

/* the contributions of particle p's face deposit (base cell (ii, jj, kk):
 * plane kk for the x, y and z faces, plane kk + 1 for the z faces) that land
 * in plane t (t < 0: all), in the order of particle.cpp:494-520 */
static inline void face_one(or_state *s, double **A, int p, int t)
{
  const double *xi = s->q[QX][OR_INT], *yi = s->q[QY][OR_INT], *zi = s->q[QZ][OR_INT];
  const double *ru = s->aux[OR_AUX_RHS_U], *rv = s->aux[OR_AUX_RHS_V], *rw = s->aux[OR_AUX_RHS_W];
  int ii = x2i(s, xi[p]), jj = y2j(s, yi[p]), kk = z2k(s, zi[p]);
  double xl = xi[p] * s->idx - (int)(xi[p] * s->idx);
  double yl = yi[p] * s->idy - (int)(yi[p] * s->idy);
  double zl = zi[p] * s->idz - (int)(zi[p] * s->idz);
  int lo = t < 0 || kk == t, hi = t < 0 || kk + 1 == t;
  if (lo) {
    A[0][G(s, ii + 1, jj, kk)] += xl * ru[p];
    A[0][G(s, ii, jj, kk)] += (1 - xl) * ru[p];
    A[1][G(s, ii, jj + 1, kk)] += yl * rv[p];
    A[1][G(s, ii, jj, kk)] += (1 - yl) * rv[p];
  }
  if (hi) A[2][G(s, ii, jj, kk + 1)] += zl * rw[p];
  if (lo) A[2][G(s, ii, jj, kk)] += (1 - zl) * rw[p];
}

/* P6: particle::part2gas_velocity (particle.cpp:494-520) into RHS_RU */
void or_p_deposit_velocity(or_state *s)
{
  double **A = T1(s, OR_RHS_RU);
  for (int c = 0; c < 3; ++c) ew_set(s, A[c], 0.);
  if (nthreads_now() == 1) {
    for (int p = 0; p < s->np; ++p) face_one(s, A, p, -1);
  } else {
    const double *zi = s->q[QZ][OR_INT];
    bucket_alloc(s);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < s->np; ++p) s->bk_plane[p] = z2k(s, zi[p]);
    bucket_planes(s);
    const int *idx = s->bk_idx, *st = s->bk_start;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < s->NZ; ++t) {
      int ia = st[t], ea = st[t + 1], ib = st[t + 1], eb = st[t + 2];  /* base t - 1, base t */
      for (int p; (p = MERGE_NEXT(ia, ea, ib, eb, idx)) >= 0;) face_one(s, A, p, t);
    }
  }
  for (int c = 0; c < 3; ++c) { fold(s, A[c]); fill(s, A[c]); }
}

/* P8/P10 gather: particle::gas2part_Temp_{int,new} (particle.cpp:402-456) */
void or_p_gather_temp(or_state *s, int stage)
{
  const double *T = T0(s, OR_T);
  double *x = s->q[QX][stage], *y = s->q[QY][stage], *z = s->q[QZ][stage];
  double *Tg = s->aux[OR_AUX_TG];
#pragma omp parallel for schedule(static)
  for (int p = 0; p < s->np; ++p) {
    int ii, jj, kk; double xl, yl, zl;
    cic(s, x[p], y[p], z[p], &ii, &jj, &kk, &xl, &yl, &zl);
    Tg[p] = xl * yl * zl * T[G(s, ii, jj, kk)] +
            (1 - xl) * yl * zl * T[G(s, ii - 1, jj, kk)] + xl * (1 - yl) * zl * T[G(s, ii, jj - 1, kk)] +
            xl * yl * (1 - zl) * T[G(s, ii, jj, kk - 1)] +
            (1 - xl) * (1 - yl) * zl * T[G(s, ii - 1, jj - 1, kk)] +
            (1 - xl) * yl * (1 - zl) * T[G(s, ii - 1, jj, kk - 1)] +
            xl * (yl - 1) * (zl - 1) * T[G(s, ii, jj - 1, kk - 1)] +
            (1 - xl) * (1 - yl) * (1 - zl) * T[G(s, ii - 1, jj - 1, kk - 1)];
  }
}

/* P8/P10: particle::Compute_RHS_Temp_{int,new} (particle.cpp:460-474) */
static void compute_rhs_temp(or_state *s, int stage)
{
  double iTth = s->p.nu * OR_PI * s->p.dp * s->p.k;
  const double *Tp = s->q[QT][stage];
#pragma omp parallel for schedule(static)
  for (int p = 0; p < s->np; ++p) s->aux[OR_AUX_RHS_T][p] = (s->aux[OR_AUX_TG][p] - Tp[p]) * iTth;
}

/* P9/P10 deposit: particle::part2gas_Temp_{int,new} (particle.cpp:582-654) */
void or_p_deposit_temp(or_state *s, int stage)
{
  double *A = T0(s, OR_RHS_PART_TEMP);
  ew_set(s, A, 0.);
  deposit_cic(s, A, s->q[QX][stage], s->q[QY][stage], s->q[QZ][stage], s->aux[OR_AUX_RHS_T]);
  fold(s, A);
  fill(s, A);
}

/* P7: particle::Send_Recv on one rank (particle.cpp:58-104, 150-311,
 * 772-846).  A particle whose x_new leaves [0, L] is wrapped (x_new, x and
 * x_np1 shift by L; x_int does not), removed by swapping with the last
 * particle, and re-appended: first the ones that left below 0, then the ones
 * that left above L.  Per-particle tau/eps travel with the particle. */
#define NREC 31
static void rec_store(const or_state *s, int p, double *r)
{
  int m = 0;
  for (int a = 0; a < 7; ++a) for (int b = 0; b < 4; ++b) r[m++] = s->q[a][b][p];
  r[m++] = (double)s->flag[p];
  r[m++] = s->aux[OR_AUX_TAU][p];
  r[m++] = s->aux[OR_AUX_EPS][p];
}
static void rec_load(or_state *s, int p, const double *r)
{
  int m = 0;
  for (int a = 0; a < 7; ++a) for (int b = 0; b < 4; ++b) s->q[a][b][p] = r[m++];
  s->flag[p] = (int)r[m++];
  s->aux[OR_AUX_TAU][p] = r[m++];
  s->aux[OR_AUX_EPS][p] = r[m++];
}
static void swap_p(or_state *s, int i, int j)
{
  for (int a = 0; a < 7; ++a)
    for (int b = 0; b < 4; ++b) { double t = s->q[a][b][i]; s->q[a][b][i] = s->q[a][b][j]; s->q[a][b][j] = t; }
  int t = s->flag[i]; s->flag[i] = s->flag[j]; s->flag[j] = t;
  for (int a = OR_AUX_TAU; a <= OR_AUX_EPS; ++a) { double u = s->aux[a][i]; s->aux[a][i] = s->aux[a][j]; s->aux[a][j] = u; }
}
static void mode_l(or_state *s, int axis, int p, double L)
{
  double *xn = s->q[axis][OR_NEW], *x = s->q[axis][OR_BASE], *x1 = s->q[axis][OR_NP1];
  if (xn[p] > L) { xn[p] -= L; x[p] -= L; x1[p] -= L; }
  else if (xn[p] < 0) { xn[p] += L; x[p] += L; x1[p] += L; }
}
static void sort_valid(or_state *s, int axis, double lo, double hi, double L)
{
  double *buf_hi = (double *)malloc(sizeof(double) * NREC * (size_t)(s->np + 1));
  double *buf_lo = (double *)malloc(sizeof(double) * NREC * (size_t)(s->np + 1));
  int nhi = 0, nlo = 0, i = 0;
  while (i < s->np) {
    double xn = s->q[axis][OR_NEW][i];
    if (xn > hi) {
      mode_l(s, axis, i, L); rec_store(s, i, buf_hi + NREC * nhi++);
      swap_p(s, i, s->np - 1); s->np--; i--;
    } else if (xn < lo) {
      mode_l(s, axis, i, L); rec_store(s, i, buf_lo + NREC * nlo++);
      swap_p(s, i, s->np - 1); s->np--; i--;
    }
    i++;
  }
  for (int m = 0; m < nlo; ++m) rec_load(s, s->np++, buf_lo + NREC * m);
  for (int m = 0; m < nhi; ++m) rec_load(s, s->np++, buf_hi + NREC * m);
  free(buf_hi); free(buf_lo);
}
void or_p_send_recv(or_state *s)
{
  sort_valid(s, QX, 0.0, s->xh, s->p.lx);
  sort_valid(s, QY, 0.0, s->yh, s->p.ly);
  sort_valid(s, QZ, 0.0, s->zh, s->p.lz);
}

/* ------------------------------------------------------------------ */
/* Poisson: poisson::FFT and poisson::CCP_FFT (poisson.cpp:391-462)    */
/* ------------------------------------------------------------------ */
static void ccp_fft(or_state *s, const double *rhs, double *out)
{
  int nx = s->nx, ny = s->ny, nz = s->nz;
  double tp = 2 * 3.141592653589793;
  double tpx = tp / s->p.lx, tpy = tp / s->p.ly, tpz = tp / s->p.lz;
  double *d = s->fft;
  long m = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++m) {
        d[2 * m + 1] = 0;
        d[2 * m] = rhs[G(s, i + s->bs, j + s->bs, k + s->bs)];
      }
  or_fft3d(d, nx, ny, nz, -1);
  double dx2 = 2. / (s->dx * s->dx), dy2 = 2. / (s->dy * s->dy), dz2 = 2. / (s->dz * s->dz);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long c = ((long)k * ny + j) * nx + i;
        if (i + j + k > 0) {
          double ii = ((i < nx / 2) ? i : i - nx) * tpx;
          double jj = ((j < ny / 2) ? j : j - ny) * tpy;
          double kk = ((k < nz / 2) ? k : k - nz) * tpz;
          double lam = (dx2) * (1 - cos(ii * s->dx)) + (dy2) * (1 - cos(jj * s->dy)) + (dz2) * (1 - cos(kk * s->dz));
          d[2 * c + 1] /= lam;
          d[2 * c] /= lam;
        }
      }
  or_fft3d(d, nx, ny, nz, +1);
  double tot = (double)(nx * ny * nz);
  m = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++m)
        out[G(s, i + s->bs, j + s->bs, k + s->bs)] = (-1) * d[2 * m] / (tot);
}

/* poisson::FFT (poisson.cpp:391-411) with C = coefficient tensor1 */
static void poisson_fft(or_state *s, double *const *C, double *P, const double *RHS)
{
  i_f2c_mean(s, s->cc, C);
  ew_sdiv(s, s->cc, -1., s->cc);
  int it = 0;
  double err = 100;
  while ((it < s->Iteration) && (err > s->p.epsilon)) {
    grad_c2f(s, s->pd, P);
    for (int c = 0; c < 3; ++c) ew_mul(s, s->pd[c], C[c]);
    div_f2c(s, s->rhs_new, s->pd);
    ew_sub(s, s->rhs_new, RHS);
    ew_mul(s, s->rhs_new, s->cc);
    ccp_fft(s, s->rhs_new, s->pd[0]);
    ew_add(s, P, s->pd[0]);
    fill(s, P);
    err = red_max_abs(s, s->pd[0]);
    it++;
  }
  s->num_iteration = it;
}

/* ------------------------------------------------------------------ */
/* grid methods                                                        */
/* ------------------------------------------------------------------ */
void or_set_rk4_count(or_state *s, int rk) { s->rk = rk; }

/* params::update (params.cpp:75-86) */
void or_params_update(or_state *s)
{
  if (s->T_cur > s->p.threshold) {
    s->Iteration = s->p.iteration2; s->A = s->p.a2;
    s->gx = s->p.gx2; s->gy = s->p.gy2; s->gz = s->p.gz2; s->I0 = s->p.i02;
  }
}

/* G14: prep tail of grid::Initialize (grid.cpp:460-486) */
void or_initialize(or_state *s)
{
  for (int c = 0; c < 3; ++c) {
    ew_copy(s, T1(s, OR_RU_INT)[c], T1(s, OR_RU)[c]);
    ew_copy(s, T1(s, OR_RU_NP1)[c], T1(s, OR_RU)[c]);
  }
  ew_copy(s, T0(s, OR_RHO_INT), T0(s, OR_RHO));
  ew_copy(s, T0(s, OR_RHO_NP1), T0(s, OR_RHO));
  for (int c = 0; c < 3; ++c) {  /* grid.cpp:463-464 */
    ew_copy(s, T1(s, OR_RV_INT)[c], T1(s, OR_RV)[c]);
    ew_copy(s, T1(s, OR_RV_NP1)[c], T1(s, OR_RV)[c]);
  }
  ew_copy(s, T0(s, OR_PS_INT), T0(s, OR_PS));  /* grid.cpp:465-466, 476 */
  ew_copy(s, T0(s, OR_PS_NP1), T0(s, OR_PS));
  i_c2f(s, T1(s, OR_PS_FACE), T0(s, OR_PS));
  s->P0_int = s->P0; s->P0_np1 = s->P0;
  i_c2f(s, T1(s, OR_RHO_FACE), T0(s, OR_RHO));
  for (int a = 0; a < 7; ++a) {
    memcpy(s->q[a][OR_INT], s->q[a][OR_BASE], sizeof(double) * (size_t)s->np);
    memcpy(s->q[a][OR_NP1], s->q[a][OR_BASE], sizeof(double) * (size_t)s->np);
  }
  ew_sdiv(s, T0(s, OR_T), s->P0_int / s->R, T0(s, OR_RHO_INT));
  or_p_gather_temp(s, OR_INT);
  compute_rhs_temp(s, OR_INT);
  or_p_deposit_temp(s, OR_INT);
}

void or_set_restart_numbers(or_state *s, double P0, double T_cur, int num_timestep)
{
  s->P0 = P0; s->T_cur = T_cur; s->num_timestep = num_timestep;
}

/* G1: grid::Update_Rho (grid.cpp:608-614) */
void or_update_rho(or_state *s)
{
  double dt = s->p.dt;
  div_f2c(s, T0(s, OR_RHS_RHO), T1(s, OR_RU_INT));
  ew_pem(s, T0(s, OR_RHO_NP1), -(dt * s->post[s->rk]), T0(s, OR_RHS_RHO));
  if (s->rk != 3) ew_lc2(s, T0(s, OR_RHO_NEW), 1, T0(s, OR_RHO), -dt * s->pre[s->rk], T0(s, OR_RHS_RHO));
  else ew_copy(s, T0(s, OR_RHO_NEW), T0(s, OR_RHO_NP1));
}

/* G2: grid::Update_P0 (grid.cpp:816-830) */
void or_update_p0(or_state *s)
{
  s->met = red_mean(s, T0(s, OR_RHS_PART_TEMP));
  if (s->p.cooling) s->dP0_dt = 0;
  else s->dP0_dt = -s->R / s->p.cv * s->met / s->Vcell;
  s->P0_np1 += s->p.dt * s->post[s->rk] * s->dP0_dt;
  if (s->rk != 3) s->P0_new = s->P0 + s->p.dt * s->pre[s->rk] * s->dP0_dt;
  else s->P0_new = s->P0_np1;
}

/* G3: grid::Update_Particle (grid.cpp:865-888) */
void or_update_particle(or_state *s)
{
  for (int c = 0; c < 3; ++c) ew_div(s, T1(s, OR_U)[c], T1(s, OR_RU_INT)[c], T1(s, OR_RHO_FACE)[c]);
  or_p_gather_velocity(s);
  update_position(s);
  update_velocity(s);
  update_temp(s);
  or_p_deposit_concentration(s);
  if (s->p.twc) or_p_deposit_velocity(s);
  else for (int c = 0; c < 3; ++c) ew_set(s, T1(s, OR_RHS_RU)[c], 0.);
  or_p_send_recv(s);
  ew_sdiv(s, T0(s, OR_T), s->P0_new / s->R, T0(s, OR_RHO_NEW));
  or_p_gather_temp(s, OR_NEW);
  compute_rhs_temp(s, OR_NEW);
  or_p_deposit_temp(s, OR_NEW);
}

/* G4: grid::Update_RU_WOP (grid.cpp:776-814) */
void or_update_ru_wop(or_state *s)
{
  double **U = T1(s, OR_U), **RUi = T1(s, OR_RU_INT), **RHS = T1(s, OR_RHS_RU);
  double **d = T1(s, OR_DUMMY), **d2 = T1(s, OR_DUMMY2), **Rf = T1(s, OR_RHO_FACE);
  double *dv = T0(s, OR_DIVERGENCE);
  double dt = s->p.dt;
  div_f2c(s, dv, U);
  grad_c2f(s, d2, dv);
  for (int c = 0; c < 3; ++c) del2(s, d[c], U[c]);
  for (int c = 0; c < 3; ++c)
    ew_lc3(s, RHS[c], -s->mp / s->Vcell, RHS[c], s->p.mu0 / 3., d2[c], s->p.mu0, d[c]);
  for (int a = 0; a < 3; ++a) {          /* convection of component a */
    i_c2f(s, d, U[a]);
    iaxis_c2f_t1(s, d2, RUi, a);
    for (int c = 0; c < 3; ++c) ew_mul(s, d[c], d2[c]);
    div_f2c(s, d2[a], d);
    ew_sub(s, RHS[a], d2[a]);
  }
  for (int c = 0; c < 3; ++c) ew_pem(s, RHS[c], s->A, RUi[c]);
  for (int c = 0; c < 3; ++c) ew_sub_s(s, Rf[c], s->p.rho0);
  ew_mulc(s, d[0], Rf[0], s->gx);
  ew_mulc(s, d[1], Rf[1], s->gy);
  ew_mulc(s, d[2], Rf[2], s->gz);
  for (int c = 0; c < 3; ++c) ew_add(s, RHS[c], d[c]);
  for (int c = 0; c < 3; ++c) {
    ew_pem(s, T1(s, OR_RU_NP1)[c], dt * s->post[s->rk], RHS[c]);
    if (s->rk != 3) ew_lc2(s, T1(s, OR_RU_NEW)[c], 1, T1(s, OR_RU)[c], dt * s->pre[s->rk], RHS[c]);
    else ew_copy(s, T1(s, OR_RU_NEW)[c], T1(s, OR_RU_NP1)[c]);
  }
}

/* G5: grid::Compute_Div_U_new (grid.cpp:832-839) */
void or_compute_div_u_new(or_state *s)
{
  double **d = T1(s, OR_DUMMY);
  ew_sdiv(s, d[0], 1., T0(s, OR_RHO_NEW));
  del2(s, d[1], d[0]);
  ew_copy(s, d[2], T0(s, OR_RHS_PART_TEMP));
  if (s->p.cooling) make_mean_zero(s, d[2]);
  ew_lc2(s, T0(s, OR_DIVERGENCE), s->p.k / s->p.cp, d[1], -s->R / (s->p.cp * s->P0_new * s->Vcell), d[2]);
}

/* G6: grid::Compute_RHS_Pois (grid.cpp:841-849) */
void or_compute_rhs_pois(or_state *s)
{
  double **Rf = T1(s, OR_RHO_FACE), **U = T1(s, OR_U);
  double *R = T0(s, OR_RHS_POIS);
  i_c2f(s, Rf, T0(s, OR_RHO_NEW));
  for (int c = 0; c < 3; ++c) ew_div(s, U[c], T1(s, OR_RU_NEW)[c], Rf[c]);
  div_f2c(s, R, U);
  ew_sub(s, R, T0(s, OR_DIVERGENCE));
  ew_mul_s(s, R, (1. / (s->p.dt * s->pre[s->rk])));
  make_mean_zero(s, R);
}

/* G7: grid::Solve_Poisson (grid.cpp:851-855) -> poisson::Solve (Solver==0) */
void or_solve_poisson(or_state *s)
{
  double **d = T1(s, OR_DUMMY);
  for (int c = 0; c < 3; ++c) ew_sdiv(s, d[c], 1., T1(s, OR_RHO_FACE)[c]);
  poisson_fft(s, d, T0(s, OR_P), T0(s, OR_RHS_POIS));
  s->total_iterations += s->num_iteration;
}

/* poisson::Solve(tensor1& C, tensor0& P, tensor0& RHS) (poisson.cpp:115-117)
 * with the caller's coefficient: C, P and RHS in the interior layout, ghosts
 * filled here; P is the initial guess and receives the result.  The
 * coefficient is held in the grid's dummy2 tensor, as grid::Solve_Poisson_Q
 * holds its coefficient in dummy (grid.cpp:666-670). */
void or_poisson_solve(or_state *s, const double *cx, const double *cy, const double *cz, double *P,
                      const double *rhs, int *iters)
{
  double **C = T1(s, OR_DUMMY2);
  const double *cin[3] = {cx, cy, cz};
  for (int c = 0; c < 3; ++c) load_int(s, C[c], cin[c]);
  load_int(s, s->scratch[0], P);
  load_int(s, s->scratch[1], rhs);
  poisson_fft(s, C, s->scratch[0], s->scratch[1]);
  store_int(s, s->scratch[0], P);
  if (iters) *iters = s->num_iteration;
}

/* G10: grid::Update_RU_WP (grid.cpp:857-863) */
void or_update_ru_wp(or_state *s)
{
  double **d = T1(s, OR_DUMMY), dt = s->p.dt;
  grad_c2f(s, d, T0(s, OR_P));
  for (int c = 0; c < 3; ++c) {
    ew_pem(s, T1(s, OR_RU_NP1)[c], -(dt * s->post[s->rk]), d[c]);
    if (s->rk != 3) ew_pem(s, T1(s, OR_RU_NEW)[c], -(dt * s->pre[s->rk]), d[c]);
    else ew_copy(s, T1(s, OR_RU_NEW)[c], T1(s, OR_RU_NP1)[c]);
  }
}

/* G11: grid::TimeAdvance_RK4 (grid.cpp:890-900) */
void or_time_advance_rk4(or_state *s)
{
  ew_copy(s, T0(s, OR_RHO_INT), T0(s, OR_RHO_NEW));
  for (int c = 0; c < 3; ++c) ew_copy(s, T1(s, OR_RU_INT)[c], T1(s, OR_RU_NEW)[c]);
  for (int c = 0; c < 3; ++c) ew_copy(s, T1(s, OR_RV_INT)[c], T1(s, OR_RV_NEW)[c]);  /* grid.cpp:894 */
  ew_copy(s, T0(s, OR_PS_INT), T0(s, OR_PS_NEW));  /* grid.cpp:897 */
  s->P0_int = s->P0_new;
  for (int a = 0; a < 7; ++a) memcpy(s->q[a][OR_INT], s->q[a][OR_NEW], sizeof(double) * (size_t)s->np);
}

/* grid::C_Source (grid.cpp:686-708) with scalar_source.h:4-8: A cos(k x),
 * A = k = 1, at the cell centre X = dx/2 + i dx (the reference passes
 * (X, Y, X, T); the function reads only its first argument) */
static void c_source(or_state *s, double *o)
{
  IN_LOOP(s) {
    long c = G(s, i, j, k);
    double X = s->dx / 2. + (i - s->bs) * s->dx;
    double A = 1., kk = 1.;
    o[c] = A * cos(kk * X);
  }
  fill(s, o);
}

/* grid::Update_Passive_Scalar (grid.cpp:711-745): flux-form transport by
 * U = RU_int/Rho_face of the scalar on the faces (I_C2F of the previous
 * substep's new value), molecular diffusion D_M del2, a source (S1_type 0:
 * C_Source minus its mean; else -Ix_F2C(U.x)), then the RK4 updates. */
void or_update_passive_scalar(or_state *s)
{
  double dt = s->p.dt;
  double *const *U = T1(s, OR_U), *const *D = T1(s, OR_DUMMY), *const *D2 = T1(s, OR_DUMMY2);
  double *const *F = T1(s, OR_PS_FACE);
  double *rhs = T0(s, OR_RHS_PS), *S1 = T0(s, OR_S1);
  for (int c = 0; c < 3; ++c) ew_div(s, U[c], T1(s, OR_RU_INT)[c], T1(s, OR_RHO_FACE)[c]);
  for (int c = 0; c < 3; ++c) {  /* tensor1::Equal_Mult(U, Passive_Scalar_face) */
    double *o = D[c];
    const double *a = U[c], *b = F[c];
#pragma omp parallel for schedule(static)
    EW_LOOP o[e] = a[e] * b[e];
  }
  div_f2c(s, D2[0], D);
  del2(s, D2[1], T0(s, OR_PS_INT));
  c_source(s, S1);
  ew_lc2(s, rhs, s->p.d_m, D2[1], -1, D2[0]);
  if (!s->p.s1_type) {
    ew_sub_s(s, S1, red_mean(s, S1));
    ew_add(s, rhs, S1);
  } else {
    i_f2c_axis(s, D[0], U[0], 0);
    ew_pem(s, rhs, -1, D[0]);
  }
  ew_pem(s, T0(s, OR_PS_NP1), dt * s->post[s->rk], rhs);
  if (s->rk != 3) ew_lc2(s, T0(s, OR_PS_NEW), 1, T0(s, OR_PS), dt * s->pre[s->rk], rhs);
  else ew_copy(s, T0(s, OR_PS_NEW), T0(s, OR_PS_NP1));
  i_c2f(s, F, T0(s, OR_PS_NEW));
}

/* ------------------------------------------------------------------ */
/* passive vector (grid.cpp:616-715)                                    */
/* ------------------------------------------------------------------ */
/* Equal_Divide(tensor0 a, double b): o = a/b (tensor0.cpp:143-147) */
static void ew_divs(const or_state *s, double *o, const double *a, double b) {
#pragma omp parallel for schedule(static)
  EW_LOOP o[e] = a[e] / b; }

/* grid::V_Source (grid.cpp:616-641) with vector_source.h:4-15: S2.x =
 * Vector_Source_x(X - dx/2, Y, Z) = 0, S2.y = Vector_Source_y(X, Y - dy/2, Z)
 * = A cos(K X) (A = K = 1, X the cell-centre x), S2.z = 0 */
static void v_source(or_state *s, double *const *S2)
{
  IN_LOOP(s) {
    long c = G(s, i, j, k);
    double X = (i - s->bs) * s->dx + s->dx / 2.;
    double A = 1., K = 1.;
    S2[0][c] = 0.;
    S2[1][c] = A * cos(K * X);
    S2[2][c] = 0.;
  }
  fill(s, S2[0]); fill(s, S2[1]); fill(s, S2[2]);
}

/* tensor0::Equal_Iy_C2F(T) called with T the tensor itself
 * (grid.cpp:684: dummy.x.Equal_Iy_C2F(dummy.x)): the loop of
 * tensor0.cpp:308-315 runs i, then j upwards, then k, and reads row j - 1
 * after it has been overwritten, so row j becomes (f_j + new_{j-1}) / 2, the
 * first interior row taking its ghost neighbour's old value. */
static void iy_c2f_in_place(or_state *s, double *f)
{
  long sy = s->NX;
#pragma omp parallel for schedule(static)
  for (int i = s->bs; i < s->NX - s->bs; ++i)
    for (int j = s->bs; j < s->NY - s->bs; ++j)
      for (int k = s->bs; k < s->NZ - s->bs; ++k) {
        long c = G(s, i, j, k);
        f[c] = (f[c] + f[c - sy]) * 0.5;
      }
  fill(s, f);
}

/* grid::Update_RV_WOQ (grid.cpp:646-692): V = RV_int/Rho_forV; RHS_RV =
 * eta/3 grad(div V) + eta del2 V - div(I_C2F(RV_int.a) I_a(U)) per component
 * a (U = RU_int/Rho_face, left by Update_Particle), plus the source (S2_type
 * 0: V_Source minus its means; else -Iy_C2F of RU_int.x/Rho_face.x on the y
 * component), then the RK4 updates. */
void or_update_rv_woq(or_state *s)
{
  double dt = s->p.dt, eta = s->p.eta;
  double **V = T1(s, OR_V), **RVi = T1(s, OR_RV_INT), **RHS = T1(s, OR_RHS_RV), **U = T1(s, OR_U);
  double **d = T1(s, OR_DUMMY), **d2 = T1(s, OR_DUMMY2), **S2 = T1(s, OR_S2);
  double *dv = T0(s, OR_DIVERGENCE);
  for (int c = 0; c < 3; ++c) ew_divs(s, V[c], RVi[c], s->p.rho_forv);
  div_f2c(s, dv, V);
  grad_c2f(s, d2, dv);
  for (int c = 0; c < 3; ++c) del2(s, d[c], V[c]);
  for (int c = 0; c < 3; ++c) ew_lc2(s, RHS[c], eta / 3., d2[c], eta, d[c]);
  for (int a = 0; a < 3; ++a) {          /* convection of component a */
    i_c2f(s, d, RVi[a]);
    iaxis_c2f_t1(s, d2, U, a);
    for (int c = 0; c < 3; ++c) ew_mul(s, d[c], d2[c]);
    div_f2c(s, d2[a], d);
    ew_sub(s, RHS[a], d2[a]);
  }
  if (!s->p.s2_type) {
    v_source(s, S2);
    for (int c = 0; c < 3; ++c) ew_sub_s(s, S2[c], red_mean(s, S2[c]));
    for (int c = 0; c < 3; ++c) ew_add(s, RHS[c], S2[c]);
  } else {
    ew_div(s, d[0], T1(s, OR_RU_INT)[0], T1(s, OR_RHO_FACE)[0]);
    iy_c2f_in_place(s, d[0]);
    ew_pem(s, RHS[1], -1, d[0]);
  }
  for (int c = 0; c < 3; ++c) {
    ew_pem(s, T1(s, OR_RV_NP1)[c], dt * s->post[s->rk], RHS[c]);
    if (s->rk != 3) ew_lc2(s, T1(s, OR_RV_NEW)[c], 1, T1(s, OR_RV)[c], dt * s->pre[s->rk], RHS[c]);
    else ew_copy(s, T1(s, OR_RV_NEW)[c], T1(s, OR_RV_NP1)[c]);
  }
}

/* grid::Compute_RHS_Pois_Q (grid.cpp:694-701) */
void or_compute_rhs_pois_q(or_state *s)
{
  double **V = T1(s, OR_V);
  double *R = T0(s, OR_RHS_POIS_Q);
  for (int c = 0; c < 3; ++c) ew_divs(s, V[c], T1(s, OR_RV_NEW)[c], s->p.rho_forv);
  div_f2c(s, R, V);
  ew_mul_s(s, R, (1. / (s->p.dt * s->pre[s->rk])));
  make_mean_zero(s, R);
}

/* grid::Solve_Poisson_Q (grid.cpp:703-707): coefficient 1/Rho_forV on every
 * face; the solver object is the gas's, so num_iteration is shared */
void or_solve_poisson_q(or_state *s)
{
  double **d = T1(s, OR_DUMMY);
  for (int c = 0; c < 3; ++c) ew_set(s, d[c], 1. / s->p.rho_forv);
  poisson_fft(s, d, T0(s, OR_Q), T0(s, OR_RHS_POIS_Q));
}

/* grid::Update_RV_WQ (grid.cpp:709-715) */
void or_update_rv_wq(or_state *s)
{
  double **d = T1(s, OR_DUMMY), dt = s->p.dt;
  grad_c2f(s, d, T0(s, OR_Q));
  for (int c = 0; c < 3; ++c) {
    ew_pem(s, T1(s, OR_RV_NP1)[c], -(dt * s->post[s->rk]), d[c]);
    if (s->rk != 3) ew_pem(s, T1(s, OR_RV_NEW)[c], -(dt * s->pre[s->rk]), d[c]);
    else ew_copy(s, T1(s, OR_RV_NEW)[c], T1(s, OR_RV_NP1)[c]);
  }
}

/* G12: tensor1::make_mean_U0 on RU_np1 (tensor1.cpp:406-411, tensor0.cpp:386-396) */
void or_make_mean_u0(or_state *s, const double a[3])
{
  for (int c = 0; c < 3; ++c) {
    double *f = T1(s, OR_RU_NP1)[c];
    double m = red_mean(s, f);
    ew_sub_s(s, f, m);
    ew_add_s(s, f, a[c]);
  }
}

/* G13: grid::TimeAdvance (grid.cpp:595-606) */
void or_time_advance(or_state *s)
{
  ew_copy(s, T0(s, OR_RHO), T0(s, OR_RHO_NP1));
  for (int c = 0; c < 3; ++c) ew_copy(s, T1(s, OR_RU)[c], T1(s, OR_RU_NP1)[c]);
  ew_copy(s, T0(s, OR_PS), T0(s, OR_PS_NP1));  /* grid.cpp:598 */
  for (int c = 0; c < 3; ++c) ew_copy(s, T1(s, OR_RV)[c], T1(s, OR_RV_NP1)[c]);  /* grid.cpp:600 */
  s->P0 = s->P0_np1;
  for (int a = 0; a < 7; ++a) memcpy(s->q[a][OR_BASE], s->q[a][OR_NP1], sizeof(double) * (size_t)s->np);
  s->T_cur += s->p.dt;
  s->num_timestep++;
}

/* simulation.cpp:57-93 with Update_Particle enabled at :64 and the passive
 * scalar when solve_for_scalar is set (vector, elongated box, statistics and
 * touch switch off). */
void or_step(or_state *s, int nsteps)
{
  double a[3] = {s->p.u0 * s->p.rho0, s->p.v0 * s->p.rho0, s->p.w0 * s->p.rho0};
  for (int n = 0; n < nsteps; ++n) {
    or_params_update(s);
    for (s->rk = 0; s->rk < 4; s->rk++) {
      or_update_rho(s);
      or_update_p0(s);
      or_update_particle(s);
      if (s->p.solve_for_scalar) or_update_passive_scalar(s);
      if (s->p.solve_for_vector) {  /* simulation.cpp:66-71 */
        or_update_rv_woq(s);
        or_compute_rhs_pois_q(s);
        or_solve_poisson_q(s);
        or_update_rv_wq(s);
      }
      or_update_ru_wop(s);
      or_compute_div_u_new(s);
      or_compute_rhs_pois(s);
      or_solve_poisson(s);
      or_update_ru_wp(s);
      or_time_advance_rk4(s);
    }
    or_make_mean_u0(s, a);
    or_time_advance(s);
  }
}

void or_get_scalars(const or_state *s, double *o)
{
  o[OR_S_P0] = s->P0; o[OR_S_P0_INT] = s->P0_int; o[OR_S_P0_NEW] = s->P0_new; o[OR_S_P0_NP1] = s->P0_np1;
  o[OR_S_T_CUR] = s->T_cur; o[OR_S_NUM_TIMESTEP] = s->num_timestep;
  o[OR_S_NUM_ITERATION] = s->num_iteration; o[OR_S_MEAN_ENERGY] = s->met; o[OR_S_A] = s->A;
  o[OR_S_I0] = s->I0; o[OR_S_ITERATION] = s->Iteration; o[OR_S_GX] = s->gx; o[OR_S_GY] = s->gy;
  o[OR_S_GZ] = s->gz; o[OR_S_DP0_DT] = s->dP0_dt; o[OR_S_TOTAL_ITERATIONS] = (double)s->total_iterations;
}

/* ------------------------------------------------------------------ */
/* primitives for closed-form pins                                     */
/* ------------------------------------------------------------------ */
static void load_int(or_state *s, double *f, const double *in)
{
  long m = 0;
  for (int k = s->bs; k < s->NZ - s->bs; ++k)
    for (int j = s->bs; j < s->NY - s->bs; ++j)
      for (int i = s->bs; i < s->NX - s->bs; ++i) f[G(s, i, j, k)] = in[m++];
  fill(s, f);
}
static void store_int(const or_state *s, const double *f, double *out)
{
  long m = 0;
  for (int k = s->bs; k < s->NZ - s->bs; ++k)
    for (int j = s->bs; j < s->NY - s->bs; ++j)
      for (int i = s->bs; i < s->NX - s->bs; ++i) out[m++] = f[G(s, i, j, k)];
}

int or_apply(or_state *s, int op, const double *a0, const double *a1, const double *a2,
             double *o0, double *o1, double *o2)
{
  double *in[3] = {s->scratch[0], s->scratch[1], s->scratch[2]};
  double *out[3] = {s->pd[0], s->pd[1], s->pd[2]};
  if (a0 && op != OR_OP_FOLD) load_int(s, in[0], a0);
  if (a1) load_int(s, in[1], a1);
  if (a2) load_int(s, in[2], a2);
  switch (op) {
  case OR_OP_DIV_F2C: div_f2c(s, out[0], in); store_int(s, out[0], o0); break;
  case OR_OP_GRAD_C2F: grad_c2f(s, out, in[0]); store_int(s, out[0], o0); store_int(s, out[1], o1); store_int(s, out[2], o2); break;
  case OR_OP_DEL2: del2(s, out[0], in[0]); store_int(s, out[0], o0); break;
  case OR_OP_I_C2F: i_c2f(s, out, in[0]); store_int(s, out[0], o0); store_int(s, out[1], o1); store_int(s, out[2], o2); break;
  case OR_OP_I_F2C: i_f2c_mean(s, out[0], in); store_int(s, out[0], o0); break;
  case OR_OP_IX_C2F: case OR_OP_IY_C2F: case OR_OP_IZ_C2F:
    i_c2f_axis(s, out[0], in[0], op - OR_OP_IX_C2F); store_int(s, out[0], o0); break;
  case OR_OP_IX_F2C: case OR_OP_IY_F2C: case OR_OP_IZ_F2C:
    i_f2c_axis(s, out[0], in[0], op - OR_OP_IX_F2C); store_int(s, out[0], o0); break;
  case OR_OP_FOLD:
    memcpy(in[0], a0, sizeof(double) * (size_t)s->n); fold(s, in[0]); store_int(s, in[0], o0); break;
  case OR_OP_CCP_FFT: ccp_fft(s, in[0], out[0]); store_int(s, out[0], o0); break;
  default: return -1;
  }
  return 0;
}

double or_sum(or_state *s, const double *interior)
{
  load_int(s, s->scratch[3], interior);
  return red_sum(s, s->scratch[3]);
}

double or_max_abs(or_state *s, const double *interior)
{
  load_int(s, s->scratch[3], interior);
  return red_max_abs(s, s->scratch[3]);
}
