// k_adjoint.hip -- hand-written adjoint of the particle path (gfx950).
//
// Objective J = Var(C) over the cells, C the cloud-in-cell concentration of
// the particles' final positions (particle.cpp:522-555 weights).  Under
// one-way coupling (no drag deposit, no heat exchange) the gas does not
// depend on the particles and dJ/dtau_p is the particle path's adjoint alone
// (k_adj_step_rev); under two-way coupling the stages are reversed one at a
// time between the gas substeps' reversal (k_adj_stage_rev, the gas sweep in
// k_adjoint_gas.hip).  The forward run tapes the gas stage fields every particle advance reads
// (RU_int, Rho_int of each substep) and checkpoints the particles' base state;
// the reverse sweep recomputes each checkpoint segment once from its
// checkpoint, keeping every step's start, with the same arithmetic as the
// advance kernel (P1-P3, P7 of particle.cpp:314-399, 58-104), then for each
// step backwards recomputes its four stage states and runs the transposed
// RK4 stage recursion
//   lam_k(s) = dt post_s lam_end + dt pre_s lam_Y(s+1),  lam_Y(s) = J_f(Y_s)^T lam_k(s),
//   lam_y0   = lam_end + sum_s lam_Y(s)
// with f(x, u) = (u, (u_g(x) - u)/tau + g): J_f^T (a, b) = (b du_g/dx / tau, a - b/tau)
// (u_g of each axis is linear along its own axis between two faces and piecewise
// constant across), and dJ/dtau_p = -sum_s b(s).(u_g(Y_s) - u_Y(s)) / tau^2.
// The periodic wraps shift positions by constants and drop out of the adjoint.
//
// With the heat exchange the state gains the particle temperature T: its
// stage right-hand side is iS (h + alpha eps) (particle.cpp:477-492), h =
// RHS_T = K (Tg(x; Rho, P0) - T) (particle.cpp:430-474, K = Nu pi Dp k, Tg
// the cloud-in-cell gather of T = P0/(R Rho)) evaluated by the heat exchange
// at the end of the previous substep, i.e. at this stage's state with the
// gas of this substep's start.  The same h is deposited (Q = D(x) h, the
// gather's weights) into the Q of the previous substep's divergence and its
// sum into the next Update_P0, so the reverse of h(Y) (k_adj_heat) takes
//   lam_h = iS lam_kT (the temperature update) + D(x)^T lam_Q + lam_(sum Q)
// and gives lam_T -= K lam_h, lam_x += K lam_h dTg/dx + h (dD/dx)^T lam_Q,
// lam_Rho -= K lam_h w T/Rho at the gather's corners and lam_P0 += K lam_h
// Tg/P0.  It runs between the gas sweep's divergence and its RK bookkeeping,
// where lam_Q of the substep and lam_Rho_new are at hand.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <stdexcept>

#include "tpx_internal.h"

namespace tpx {

namespace {

constexpr int kBlock = 256;

struct Ix {
  int mx, my, mz, nx;
  long nxy;
  __device__ __forceinline__ long operator()(int i, int j, int k) const {
    return (long)(k & mz) * nxy + (long)(j & my) * nx + (i & mx);
  }
};

__device__ __forceinline__ Ix make_ix(const Geo &g) {
  Ix r;
  r.mx = g.nx - 1; r.my = g.ny - 1; r.mz = g.nz - 1; r.nx = g.nx; r.nxy = g.nxy;
  return r;
}

// One substep's taped gas: RU_int (3 components) then Rho_int, n values each.
struct Gas {
  const double *f;
  long n;
};

// u_g at x (gas2part_velocity, particle.cpp:340-360, as in the advance
// kernel) and its derivative along each component's own axis
__device__ __forceinline__ void gather(const Geo &g, const Ix &ix, Gas gs, const double x[3], double ug[3],
                                       double dug[3]) {
  const double *r0 = gs.f, *r1 = gs.f + gs.n, *r2 = gs.f + 2 * gs.n, *rho = gs.f + 3 * gs.n;
  int ii = (int)(x[0] * g.idx), jj = (int)(x[1] * g.idy), kk = (int)(x[2] * g.idz);
  double xl = (x[0] * g.idx - ((int)(x[0] * g.idx)));
  double yl = (x[1] * g.idy - ((int)(x[1] * g.idy)));
  double zl = (x[2] * g.idz - ((int)(x[2] * g.idz)));
  long c000 = ix(ii, jj, kk);
  double ux1 = r0[ix(ii + 1, jj, kk)] / ((rho[ix(ii + 1, jj, kk)] + rho[c000]) * 0.5);
  double ux0 = r0[c000] / ((rho[c000] + rho[ix(ii - 1, jj, kk)]) * 0.5);
  double uy1 = r1[ix(ii, jj + 1, kk)] / ((rho[ix(ii, jj + 1, kk)] + rho[c000]) * 0.5);
  double uy0 = r1[c000] / ((rho[c000] + rho[ix(ii, jj - 1, kk)]) * 0.5);
  double uz1 = r2[ix(ii, jj, kk + 1)] / ((rho[ix(ii, jj, kk + 1)] + rho[c000]) * 0.5);
  double uz0 = r2[c000] / ((rho[c000] + rho[ix(ii, jj, kk - 1)]) * 0.5);
  ug[0] = xl * ux1 + (1 - xl) * ux0;
  ug[1] = yl * uy1 + (1 - yl) * uy0;
  ug[2] = zl * uz1 + (1 - zl) * uz0;
  dug[0] = (ux1 - ux0) * g.idx;
  dug[1] = (uy1 - uy0) * g.idy;
  dug[2] = (uz1 - uz0) * g.idz;
}

// The cloud-in-cell corners of x with the temperature/concentration weights
// (particle.cpp:402-456: ii = round(x idx), xl = x idx - ii + 1/2); corner
// (ii - 1 + ox, jj - 1 + oy, kk - 1 + oz) is index ox + 2 oy + 4 oz.
struct Cic {
  long c[8];
  double xl, yl, zl;
  __device__ __forceinline__ Cic(const Geo &g, const Ix &ix, double x, double y, double z) {
    const double fx = x * g.idx, fy = y * g.idy, fz = z * g.idz;
    const int ii = (int)(round(fx)), jj = (int)(round(fy)), kk = (int)(round(fz));
    xl = fx - ii + 0.5; yl = fy - jj + 0.5; zl = fz - kk + 0.5;
#pragma unroll
    for (int e = 0; e < 8; ++e) c[e] = ix(ii - 1 + (e & 1), jj - 1 + ((e >> 1) & 1), kk - 1 + (e >> 2));
  }
  __device__ __forceinline__ double w(int e) const {
    return ((e & 1) ? xl : 1 - xl) * ((e & 2) ? yl : 1 - yl) * ((e & 4) ? zl : 1 - zl);
  }
  // the interpolation of corner values v as nested lerps along x, y, z (the
  // heat exchange's order), and its derivatives along x, y, z (per cell width)
  __device__ __forceinline__ double lerp(const double v[8], double d[3]) const {
    double lx[4], dx[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      lx[r] = fma(xl, v[2 * r + 1] - v[2 * r], v[2 * r]);
      dx[r] = v[2 * r + 1] - v[2 * r];
    }
    const double t0 = fma(yl, lx[1] - lx[0], lx[0]), t1 = fma(yl, lx[3] - lx[2], lx[2]);
    const double e0 = fma(yl, dx[1] - dx[0], dx[0]), e1 = fma(yl, dx[3] - dx[2], dx[2]);
    const double f0 = lx[1] - lx[0], f1 = lx[3] - lx[2];
    d[0] = fma(zl, e1 - e0, e0);
    d[1] = fma(zl, f1 - f0, f0);
    d[2] = t1 - t0;
    return fma(zl, t1 - t0, t0);
  }
};

// RHS_T = K (Tg - T) at x, with T = (P0/R)/Rho on the cells, as the heat
// exchange forms it (k_particle.hip)
__device__ __forceinline__ double heat_rhs(const Cic &w, const double *rho, double at, double Tp, double K) {
  double v[8], d[3];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = at / rho[w.c[e]];
  return (w.lerp(v, d) - Tp) * K;
}

struct HeatParams {
  const double *p0;       // P0_int of the step's four substeps (device)
  const double *eps;      // per particle, or null
  double epsp;            // scalar eps when eps is null
  double iS, K, R, Dp, I0;
  int on;
};

struct StepArgs {
  Geo g;
  const double *gas;   // 4 substeps x 4 fields x n
  const double *yin[7];  // base x, y, z, u, v, w (and T) at the step's start
  double *y[7];          // the step's result (may be yin)
  double *stage;       // non-null: the 4 stage states, 4 x nq arrays of np
  const double *tau;   // per particle, or null
  double tp;           // scalar tau when tau is null
  double c_pre[4], c_post[4];
  double g3[3], L[3];
  HeatParams h;
  int nq;
  long np, ncell;
};

// One whole RK4 step of the particles' positions and velocities with the
// advance kernel's arithmetic (update_position, update_velocity, the wrap),
// one particle per thread; optionally stores the int state of each substep.
struct k_adj_step_fwd {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const StepArgs &a) {
    const long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (p >= a.np) return;
    const Geo &g = a.g;
    Ix ix = make_ix(g);
    double b[7], vi[7], v1[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) b[q] = vi[q] = v1[q] = q < a.nq ? a.yin[q][p] : 0.0;
    const double itp = a.tau ? 1. / a.tau[p] : 1. / a.tp;
    double alpha = 0.0;
    if (a.h.on) {
      const double epsp = a.h.eps ? a.h.eps[p] : a.h.epsp;
      alpha = 0.25 * epsp * TPX_PI * a.h.Dp * a.h.Dp * a.h.I0;
    }
    for (int rk = 0; rk < 4; ++rk) {
      if (a.stage)
#pragma unroll
        for (int q = 0; q < 7; ++q)
          if (q < a.nq) a.stage[(rk * a.nq + q) * a.np + p] = vi[q];
      if (a.h.on) {  // update_Temp with the RHS_T of this stage's state
        const double *rho = a.gas + (long)rk * 4 * a.ncell + 3 * a.ncell;
        const double rt = heat_rhs(Cic(g, ix, vi[0], vi[1], vi[2]), rho, a.h.p0[rk] / a.h.R, vi[6], a.h.K);
        const double n1t = v1[6] + a.c_post[rk] * a.h.iS * (rt + alpha);
        vi[6] = rk == 3 ? n1t : b[6] + a.c_pre[rk] * a.h.iS * (rt + alpha);
        v1[6] = n1t;
      }
      double ug[3], dug[3];
      gather(g, ix, Gas{a.gas + (long)rk * 4 * a.ncell, a.ncell}, vi, ug, dug);
      double nv[6], n1[6];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double ui = vi[3 + d];
        n1[d] = v1[d] + a.c_post[rk] * ui;
        nv[d] = rk == 3 ? n1[d] : b[d] + a.c_pre[rk] * ui;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double ui = vi[3 + d];
        double rhs = (ug[d] - ui) * itp;
        n1[3 + d] = v1[3 + d] + a.c_post[rk] * (rhs + a.g3[d]);
        nv[3 + d] = rk == 3 ? n1[3 + d] : b[3 + d] + a.c_pre[rk] * (rhs + a.g3[d]);
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (nv[d] > a.L[d]) {
          nv[d] -= a.L[d]; b[d] -= a.L[d]; n1[d] -= a.L[d];
        } else if (nv[d] < 0) {
          nv[d] += a.L[d]; b[d] += a.L[d]; n1[d] += a.L[d];
        }
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        v1[q] = n1[q];
        vi[q] = nv[q];
      }
    }
#pragma unroll
    for (int q = 0; q < 7; ++q)
      if (q < a.nq) a.y[q][p] = v1[q];
  }
};

struct AdjArgs {
  Geo g;
  const double *gas;    // the step's 4 substeps
  const double *stage;  // 24 arrays of np: the int state of each substep
  double *lam[6];       // adjoint of the step's result in, of its start out
  double *dtau;         // dJ/dtau accumulated
  const double *tau;
  double tp;
  double c_pre[4], c_post[4];
  long np, ncell;
};

// The transpose of the gather (pull form): lug, the adjoint of u_g at each
// particle (3 x np), goes to the two faces each component was interpolated
// from, with the gather's weights.  One thread per cell sums the x, y and z
// faces at its index: from the particles whose gather cell (the floor of
// x idx, y idy, z idz) is this cell, weight 1 - fraction, then from those of
// the cell below along that axis, weight fraction; bins in cell order, each
// cell's particles by index, so the sums do not depend on the run.
struct k_adj_gather_pull {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const double *const Y, const double *lug,
                                             const int32_t *start, const int32_t *perm, double *lup, long np) {
    const long c = (long)xcd_block(blockIdx.x, gridDim.x) * block_dim_x() + threadIdx.x;  // neighbour bins in this XCD's L2
    if (c >= g.n) return;
    Ix ix = make_ix(g);
    const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / g.nxy);
    const double id[3] = {g.idx, g.idy, g.idz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double *pos = Y + (long)d * np, *l = lug + (long)d * np;
      double acc = 0.0;
      for (int32_t q = start[c]; q < start[c + 1]; ++q) {
        const int32_t p = perm[q];
        const double f = pos[p] * id[d];
        acc += (1 - (f - ((int)f))) * l[p];
      }
      const long cm = ix(i - (d == 0), j - (d == 1), k - (d == 2));
      for (int32_t q = start[cm]; q < start[cm + 1]; ++q) {
        const int32_t p = perm[q];
        const double f = pos[p] * id[d];
        acc += (f - ((int)f)) * l[p];
      }
      lup[(long)d * g.n + c] = acc;
    }
  }
};

// Keys and counts of the bins of positions (x, y, z rows of Y): the gather's
// cell (floor), or with cic set the cloud-in-cell base corner (round - 1)
struct k_adj_bin_keys {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const double *Y, int cic, int32_t *key, int32_t *cnt,
                                             long np) {
    const long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (p >= np) return;
    Ix ix = make_ix(g);
    const double fx = Y[p] * g.idx, fy = Y[np + p] * g.idy, fz = Y[2 * np + p] * g.idz;
    int ii, jj, kk;
    if (cic) {
      ii = (int)(round(fx)) - 1; jj = (int)(round(fy)) - 1; kk = (int)(round(fz)) - 1;
    } else {
      ii = (int)fx; jj = (int)fy; kk = (int)fz;
    }
    const int32_t kc = (int32_t)ix(ii, jj, kk);
    key[p] = kc;
    atomicAdd(cnt + kc, 1);
  }
};

// The transposed RK4 step for one particle per thread.
struct k_adj_step_rev {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const AdjArgs &a) {
    const long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (p >= a.np) return;
    const Geo &g = a.g;
    Ix ix = make_ix(g);
    const double tau = a.tau ? a.tau[p] : a.tp;
    const double itp = 1. / tau;
    double lend[6], lb[6], lnext[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      lend[q] = a.lam[q][p];
      lb[q] = lend[q];  // the np1 accumulator starts as the base state
      lnext[q] = 0.0;
    }
    double dtau = 0.0;
    for (int rk = 3; rk >= 0; --rk) {
      double Y[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) Y[q] = a.stage[(rk * 6 + q) * a.np + p];
      double lk[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) lk[q] = a.c_post[rk] * lend[q] + (rk < 3 ? a.c_pre[rk] * lnext[q] : 0.0);
      if (rk < 3)
#pragma unroll
        for (int q = 0; q < 6; ++q) lb[q] += lnext[q];  // new = base + dt pre k
      double ug[3], dug[3];
      gather(g, ix, Gas{a.gas + (long)rk * 4 * a.ncell, a.ncell}, Y, ug, dug);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        lnext[d] = lk[3 + d] * dug[d] * itp;       // through u_g(x)
        lnext[3 + d] = lk[d] - lk[3 + d] * itp;     // k_x = u, and -u/tau
        dtau -= lk[3 + d] * (ug[d] - Y[3 + d]) * itp * itp;
      }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) a.lam[q][p] = lb[q] + lnext[q];  // the first int state is the base
    a.dtau[p] += dtau;
  }
};

// One stage of the transposed RK4 step, for the sweep that reverses the gas
// between the stages.  Under two-way coupling the stage's drag deposit
// S = D(RHS_u) (particle.cpp:494-520, the gather's weights) feeds the gas
// momentum right-hand side as a_drag S, so the gas sweep's lam_RHS of this
// substep returns to the particle through the gather of lam_S = a_drag lam_RHS:
// it adds to the adjoint of RHS_u, and its derivative along each axis times
// RHS_u to the adjoint of the position.
struct StageArgs {
  Geo g;
  const double *gas;     // this substep's taped gas (RU_int 3, Rho_int)
  const double *stage;   // the state entering this substep: nq arrays of np
  const double *lend;    // nq x np: adjoint of the step's result
  double *lb, *lnext;    // nq x np each: base accumulator; adjoint of the next stage's state in, of this one's out
  double *dtau;
  double *lug;           // out, 3 x np: the adjoint of u_g at each particle (for the gather's transpose), or null
  const double *ls[3];   // lam_RHS of this substep's momentum update (3 face fields), or null
  double a_drag;         // -mp/Vcell
  const double *tau;
  double tp;
  double c_pre, c_post;  // dt pre, dt post of this substep
  int last;              // the fourth substep
  // heat exchange (nq = 7): the temperature row
  double *lh;            // out: iS lam_kT, the adjoint of this stage's RHS_T through the update
  double *deps;          // dJ/deps accumulated
  double alpha_eps;      // d alpha / d eps = pi Dp^2 I0 / 4
  double iS;
  int nq;
  long np, ncell;
};

struct k_adj_stage_rev {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const StageArgs &a) {
    const long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (p >= a.np) return;
    const Geo &g = a.g;
    Ix ix = make_ix(g);
    const double itp = 1. / (a.tau ? a.tau[p] : a.tp);
    double Y[6], lk[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      if (q == a.nq) break;
      if (q < 6) Y[q] = a.stage[q * a.np + p];
      const double le = a.lend[q * a.np + p];
      if (a.last) {
        lk[q] = a.c_post * le;
        a.lb[q * a.np + p] = le;  // the np1 accumulator starts as the base state
      } else {
        const double ln = a.lnext[q * a.np + p];
        lk[q] = a.c_post * le + a.c_pre * ln;
        a.lb[q * a.np + p] += ln;  // new = base + dt pre k
      }
    }
    double ug[3], dug[3];
    gather(g, ix, Gas{a.gas, a.ncell}, Y, ug, dug);
    double ls[3] = {0.0, 0.0, 0.0}, dls[3] = {0.0, 0.0, 0.0};
    if (a.ls[0]) {  // lam_S at the particle with the gather's weights, and its derivative
      int ii = (int)(Y[0] * g.idx), jj = (int)(Y[1] * g.idy), kk = (int)(Y[2] * g.idz);
      const double fr[3] = {Y[0] * g.idx - ((int)(Y[0] * g.idx)), Y[1] * g.idy - ((int)(Y[1] * g.idy)),
                            Y[2] * g.idz - ((int)(Y[2] * g.idz))};
      const long c0 = ix(ii, jj, kk);
      const long c1[3] = {ix(ii + 1, jj, kk), ix(ii, jj + 1, kk), ix(ii, jj, kk + 1)};
      const double id[3] = {g.idx, g.idy, g.idz};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double l1 = a.a_drag * a.ls[d][c1[d]], l0 = a.a_drag * a.ls[d][c0];
        ls[d] = fr[d] * l1 + (1 - fr[d]) * l0;
        dls[d] = (l1 - l0) * id[d];
      }
    }
    double ln[6], dtau = 0.0, lug[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double rhs = (ug[d] - Y[3 + d]) * itp;
      const double lr = lk[3 + d] + ls[d];  // adjoint of RHS_u
      ln[d] = lr * itp * dug[d] + rhs * dls[d];
      ln[3 + d] = lk[d] - lr * itp;
      dtau -= lr * (ug[d] - Y[3 + d]) * itp * itp;
      lug[d] = lr * itp;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) a.lnext[q * a.np + p] = ln[q];
    a.dtau[p] += dtau;
    if (a.nq == 7) {  // k_T = iS (RHS_T + alpha eps): T enters only through RHS_T (k_adj_heat)
      a.lnext[6 * a.np + p] = 0.0;
      a.lh[p] = a.iS * lk[6];
      a.deps[p] += a.iS * lk[6] * a.alpha_eps;
    }
    if (a.lug)
#pragma unroll
      for (int d = 0; d < 3; ++d) a.lug[d * a.np + p] = lug[d];
  }
};

// The reverse of one heat exchange h(Y) = RHS_T at the state Y (with its
// deposit Q = D(x) h and sum Q): see the header.  lam (nq x np, T row 6)
// is the adjoint of Y, accumulated; lam_P0 of the exchange goes to the block
// partials.
struct HeatRevArgs {
  Geo g;
  const double *Y;       // nq x np: x, y, z at rows 0-2, T at row 6
  const double *rho;     // Rho of the exchange
  const double *p0;      // P0 of the exchange (device)
  const double *lq;      // adjoint of the deposit Q, or null
  const double *lsq;     // adjoint of sum Q (device scalar)
  const double *lh;      // per particle: adjoint of RHS_T through the temperature update
  double *lam;
  double *gt;            // out, per particle: the adjoint of the gathered gas temperature
  double *partials;
  double K, R;
  long np;
};

struct k_adj_heat {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const HeatRevArgs &a) {
    const long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    double lp0 = 0.0;
    if (p < a.np) {
      const Geo &g = a.g;
      Ix ix = make_ix(g);
      const double x = a.Y[p], y = a.Y[a.np + p], z = a.Y[2 * a.np + p], Tp = a.Y[6 * a.np + p];
      const Cic w(g, ix, x, y, z);
      const double P0 = a.p0[0], at = P0 / a.R;
      double T[8], dT[3];
#pragma unroll
      for (int e = 0; e < 8; ++e) T[e] = at / a.rho[w.c[e]];
      const double Tg = w.lerp(T, dT);
      const double h = (Tg - Tp) * a.K;
      double lqi = 0.0, dq[3] = {0.0, 0.0, 0.0};
      if (a.lq) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = a.lq[w.c[e]];
        lqi = w.lerp(v, dq);
      }
      const double lh = a.lh[p] + a.lsq[0] + lqi;
      const double gT = a.K * lh;  // adjoint of Tg
      const double id[3] = {g.idx, g.idy, g.idz};
#pragma unroll
      for (int d = 0; d < 3; ++d) a.lam[d * a.np + p] += (gT * dT[d] + h * dq[d]) * id[d];
      a.lam[6 * a.np + p] -= gT;
      a.gt[p] = gT;
      lp0 = gT * Tg / P0;
    }
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = lp0;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.partials[blockIdx.x] = sh[0];
  }
};

// The exchange's corner gather transposed into lam_Rho (pull form): the
// gathered T = (P0/R)/Rho at the eight corners, so each corner cell e of a
// particle gets -gT w_e T_e / Rho_e.  One thread per cell sums, for the eight
// corner offsets in turn, the particles whose cloud-in-cell base corner sits
// that offset below it (bins in cell order, each cell's by index).
struct k_adj_heat_pull {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const double *Y, const double *gt, const double *rho,
                                             const double *p0, double R, const int32_t *start, const int32_t *perm,
                                             double *lrho, long np) {
    const long c = (long)xcd_block(blockIdx.x, gridDim.x) * block_dim_x() + threadIdx.x;  // neighbour bins in this XCD's L2
    if (c >= g.n) return;
    Ix ix = make_ix(g);
    const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / g.nxy);
    const double rc = rho[c], Tc = (p0[0] / R) / rc;
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long cb = ix(i - (e & 1), j - ((e >> 1) & 1), k - (e >> 2));
      for (int32_t q = start[cb]; q < start[cb + 1]; ++q) {
        const int32_t p = perm[q];
        const Cic w(g, ix, Y[p], Y[np + p], Y[2 * np + p]);
        acc += -gt[p] * w.w(e) * Tc / rc;
      }
    }
    lrho[c] += acc;
  }
};

// acc[0] += sum of partials (one block)
struct k_adj_acc1 {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(const double *partials, long nparts, double *acc) {
    __shared__ double sh[1024];
    double v = 0.0;
    for (long q = threadIdx.x; q < nparts; q += 1024) v += partials[q];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += sh[0];
  }
};

// the adjoint of the step's start: base accumulator + the first stage's state
struct k_adj_stage_end {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const double *lb, const double *lnext, double *lam, long n6) {
    const long i = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (i < n6) lam[i] = lb[i] + lnext[i];
  }
};

// dJ/dC = 2 (C - mean C) / ncell, and the block partials of C^2
struct k_var_terms {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const double *C, long n, double *partials) {
    __shared__ double sh[kBlock];
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    const double v = c < n ? C[c] : 0.0;
    sh[threadIdx.x] = v * v;
    __syncthreads();
    for (int s = block_dim_x() / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
  }
};

struct k_dj_dc {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const double *C, long n, const double *scal, double *out) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (c >= n) return;
    const double mean = scal[SC_SUM_C] / (double)n;
    out[c] = 2.0 * (C[c] - mean) / (double)n;
  }
};

// lam_x of each particle: the concentration weights' position derivatives
// (particle.cpp:522-555 weights) against dJ/dC; lam_u = 0
struct k_conc_adj {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const double *x, const double *y, const double *z,
                                             const int32_t *cum, const double *djdc, double *l0, double *l1,
                                             double *l2, double *l3, double *l4, double *l5, double *l6,
                                             long np) {
    const long i = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (i >= np) return;
    const long p = cum[i];  // the particle at i, in the recording's order
    Ix ix = make_ix(g);
    const double xp = x[i], yp = y[i], zp = z[i];
    int ii = (int)(round(xp * g.idx)), jj = (int)(round(yp * g.idy)), kk = (int)(round(zp * g.idz));
    double xl = xp * g.idx - ii + 0.5, yl = yp * g.idy - jj + 0.5, zl = zp * g.idz - kk + 0.5;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int oz = 0; oz < 2; ++oz)
      for (int oy = 0; oy < 2; ++oy)
        for (int ox = 0; ox < 2; ++ox) {
          // corner (ii - 1 + ox, jj - 1 + oy, kk - 1 + oz) has weight wx wy wz
          const double wx = ox ? xl : 1 - xl, wy = oy ? yl : 1 - yl, wz = oz ? zl : 1 - zl;
          const double sx = ox ? g.idx : -g.idx, sy = oy ? g.idy : -g.idy, sz = oz ? g.idz : -g.idz;
          const double d = djdc[ix(ii - 1 + ox, jj - 1 + oy, kk - 1 + oz)];
          gx += d * sx * wy * wz;
          gy += d * wx * sy * wz;
          gz += d * wx * wy * sz;
        }
    l0[p] = gx; l1[p] = gy; l2[p] = gz;
    l3[p] = 0.0; l4[p] = 0.0; l5[p] = 0.0;
    if (l6) l6[p] = 0.0;
  }
};

// The particle order of a recorded run: cum[i] is the position at the start
// of the recording of the particle now at i.  A re-sort of the particle arrays
// (dst[q] = src[perm[q]]) is applied to cum too; checkpoints, tau and the
// sweep's arrays are kept in the recording's order.
struct k_adj_gather_i {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const int32_t *src, const int32_t *perm, int32_t *dst, long n) {
    const long q = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (q < n) dst[q] = src[perm[q]];
  }
};

// dst[q][cum[i]] = src[q][i] for the 6 (or 7, s6 set) state arrays
struct k_adj_scatter6 {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const double *s0, const double *s1, const double *s2, const double *s3,
                                             const double *s4, const double *s5, const double *s6,
                                             const int32_t *cum, double *dst, long n) {
    const long i = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (i >= n) return;
    const long p = cum[i];
    dst[p] = s0[i];
    dst[n + p] = s1[i];
    dst[2 * n + p] = s2[i];
    dst[3 * n + p] = s3[i];
    dst[4 * n + p] = s4[i];
    dst[5 * n + p] = s5[i];
    if (s6) dst[6 * n + p] = s6[i];
  }
};

unsigned grid1(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void hip_check_adj(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

HeatParams heat_params(const tpx_ctx *c, int step) {
  const AdjointTape &t = c->adj;
  HeatParams h{};
  h.on = t.heat ? 1 : 0;
  if (!t.heat) return h;
  h.p0 = t.p0t + (long)step * 4;
  h.eps = t.eps0;
  h.epsp = c->p.epsilonp;
  h.iS = 1. / (c->mp * c->p.cvp);
  h.K = c->p.nu * TPX_PI * c->p.dp * c->p.k;
  h.R = c->R;
  h.Dp = c->p.dp;
  h.I0 = t.I0[(size_t)step];
  return h;
}

StepArgs step_args(tpx_ctx *c, int step, const double *const yin[7], double *const y[7], double *stage) {
  Particles &pt = c->pt;
  AdjointTape &t = c->adj;
  StepArgs a{};
  a.g = c->g;
  a.gas = t.gas_of(step, c->g.n);
  a.nq = t.nq;
  for (int q = 0; q < t.nq; ++q) {
    a.yin[q] = yin[q];
    a.y[q] = y[q];
  }
  a.h = heat_params(c, step);
  a.stage = stage;
  a.tau = pt.has_tau ? t.tau0 : nullptr;
  a.tp = c->Tp;
  for (int r = 0; r < 4; ++r) {
    a.c_pre[r] = c->p.dt * c->pre[r];
    a.c_post[r] = c->p.dt * c->post[r];
  }
  for (int d = 0; d < 3; ++d) a.g3[d] = t.g3[(size_t)step * 3 + d];
  a.L[0] = c->p.lx; a.L[1] = c->p.ly; a.L[2] = c->p.lz;
  a.np = pt.n;
  a.ncell = c->g.n;
  return a;
}

}  // namespace

// copies this substep's RU_int and Rho_int into the tape
void adjoint_tape_substep(tpx_ctx *c, int rk, const double *const ru_int[3], const double *rho_int) {
  AdjointTape &t = c->adj;
  if (t.step < 0) return;
  if (t.step >= t.nsteps) throw std::logic_error("adjoint tape overrun");
  const long n = c->g.n;
  if (!t.segmented) {  // the segment mode tapes the gas when it replays a segment
    double *dst = t.gas_of(t.step, n) + (long)rk * 4 * n;
    for (int d = 0; d < 3; ++d) launch_copy(c, dst + d * n, ru_int[d], n);
    launch_copy(c, dst + 3 * n, rho_int, n);
  }
  if (t.heat) launch_copy(c, t.p0t + (long)t.step * 4 + rk, c->scal + SC_P0_INT, 1);
  if (rk == 0)
    for (int d = 0; d < 3; ++d) t.g3[(size_t)t.step * 3 + d] = (double)c->p.particle_gravity * c->gv[d];
}

void adjoint_checkpoint(tpx_ctx *c, int slot) {
  Particles &pt = c->pt;
  const int b = pt.st_base;
  const int nq = c->adj.nq;
  tpx_launch<k_adj_scatter6>(grid1(pt.n), kBlock, 0, c->stream, (const double *)pt.q[0][b], (const double *)pt.q[1][b],
                             (const double *)pt.q[2][b], (const double *)pt.q[3][b], (const double *)pt.q[4][b],
                             (const double *)pt.q[5][b], nq == 7 ? (const double *)pt.q[6][b] : nullptr,
                             (const int32_t *)c->adj.cum, c->adj.ckpt + (long)slot * nq * pt.n, (long)pt.n);
  TPX_LAUNCH_CHECK();
}

// a re-sort of the particle arrays by pt.perm_int is about to happen
void adjoint_track_permutation(tpx_ctx *c) {
  AdjointTape &t = c->adj;
  const long np = c->pt.n;
  tpx_launch<k_adj_gather_i>(grid1(np), kBlock, 0, c->stream, (const int32_t *)t.cum, (const int32_t *)c->pt.perm_int,
                             t.cum_tmp, np);
  TPX_LAUNCH_CHECK();
  std::swap(t.cum, t.cum_tmp);
}

double adjoint_objective(tpx_ctx *c) {
  const long n = c->g.n;
  // bins of the final base positions (also the int state at a step boundary)
  launch_sort_stage(c, TPX_BASE, true);
  launch_concentration(c);
  const unsigned nb = grid1(n);
  tpx_launch<k_var_terms>(nb, kBlock, 0, c->stream, (const double *)c->C, n, c->partials);
  TPX_LAUNCH_CHECK();
  launch_finish_sum(c, (int)nb, 0, SC_RED);
  tpx_launch<k_dj_dc>(nb, kBlock, 0, c->stream, (const double *)c->C, n, (const double *)c->scal, c->adj.djdc);
  TPX_LAUNCH_CHECK();
  hip_check_adj(hipStreamSynchronize(c->stream));
  double s[2];
  hip_check_adj(hipMemcpy(&s[0], c->scal + SC_SUM_C, sizeof(double), hipMemcpyDeviceToHost));
  hip_check_adj(hipMemcpy(&s[1], c->scal + SC_RED, sizeof(double), hipMemcpyDeviceToHost));
  const double mean = s[0] / (double)n;
  return s[1] / (double)n - mean * mean;
}

// Bins of the positions of the state Y (rows x, y, z of nq x np) into the
// tape's bin arrays: by the gather's cell, or (cic) the cloud-in-cell base corner.
void bin_positions(tpx_ctx *c, const double *Y, bool cic) {
  AdjointTape &t = c->adj;
  const long np = c->pt.n;
  launch_zero_counts(c, t.bstart);
  tpx_launch<k_adj_bin_keys>(grid1(np), kBlock, 0, c->stream, c->g, Y, (int)cic, t.bkey, t.bstart, np);
  TPX_LAUNCH_CHECK();
  launch_bin_by_key(c, t.bkey, t.bperm, t.binv, t.bstart);
}

// The reverse of one heat exchange at the state Y (nq x np, the recording's
// order) into lam; its lam_P0 is added to scal[slot].
void heat_reverse(tpx_ctx *c, const double *Y, const GasSweep::HeatMid &m, const double *p0, double *lam, int slot) {
  AdjointTape &t = c->adj;
  HeatRevArgs a{};
  a.g = c->g;
  a.Y = Y;
  a.rho = m.rho;
  a.p0 = p0;
  a.lq = m.lq;
  a.lsq = c->scal + SC_ADJ_LSQ;
  a.lh = t.lh;
  a.lam = lam;
  a.gt = t.gt;
  a.partials = t.hpart;
  a.K = c->p.nu * TPX_PI * c->p.dp * c->p.k;
  a.R = c->R;
  a.np = c->pt.n;
  const unsigned nb = grid1(a.np);
  tpx_launch<k_adj_heat>(nb, kBlock, 0, c->stream, a);
  TPX_LAUNCH_CHECK();
  tpx_launch<k_adj_acc1>(1, 1024, 0, c->stream, (const double *)t.hpart, (long)nb, c->scal + slot);
  TPX_LAUNCH_CHECK();
  bin_positions(c, Y, true);
  tpx_launch<k_adj_heat_pull>(grid1d(c->g.n, kBlock), kBlock, 0, c->stream, c->g, Y, (const double *)t.gt, m.rho, p0,
                              c->R, (const int32_t *)t.bstart, (const int32_t *)t.bperm, m.lrho, a.np);
  TPX_LAUNCH_CHECK();
}

void adjoint_reverse(tpx_ctx *c, const GasSweep *gas) {
  Particles &pt = c->pt;
  AdjointTape &t = c->adj;
  const long np = pt.n;
  const int nq = t.nq;
  if (t.heat && !gas) throw std::logic_error("adjoint: the heat exchange needs the gas sweep");
  double *lam[7] = {}, *y[7] = {};
  for (int q = 0; q < nq; ++q) {
    lam[q] = t.lam + q * np;
    y[q] = t.y + q * np;
  }
  // objective: lam_x from the concentration weights at the final positions
  tpx_launch<k_conc_adj>(grid1(np), kBlock, 0, c->stream, c->g, (const double *)pt.q[0][pt.st_base],
                         (const double *)pt.q[1][pt.st_base], (const double *)pt.q[2][pt.st_base],
                         (const int32_t *)t.cum, (const double *)t.djdc, lam[0], lam[1], lam[2], lam[3], lam[4],
                         lam[5], lam[6], np);
  TPX_LAUNCH_CHECK();
  launch_fill(c, t.dtau, 0.0, np);
  if (t.heat) {
    launch_fill(c, t.deps, 0.0, np);
    launch_fill(c, t.lh, 0.0, np);
  }
  // the start of every step of the checkpoint segment being reversed,
  // recomputed once per segment from its checkpoint
  if (gas && !t.lb) {
    if (hipMalloc((void **)&t.lb, sizeof(double) * nq * (size_t)np) != hipSuccess ||
        hipMalloc((void **)&t.lnext, sizeof(double) * nq * (size_t)np) != hipSuccess)
      throw std::bad_alloc();
  }
  if (gas && !t.lug) {  // the pull-form transposes: per-particle adjoints and position bins
    if (hipMalloc((void **)&t.lug, sizeof(double) * 3 * (size_t)np) != hipSuccess ||
        hipMalloc((void **)&t.gt, sizeof(double) * (size_t)np) != hipSuccess ||
        hipMalloc((void **)&t.bkey, sizeof(int32_t) * (size_t)np) != hipSuccess ||
        hipMalloc((void **)&t.bperm, sizeof(int32_t) * (size_t)np) != hipSuccess ||
        hipMalloc((void **)&t.binv, sizeof(int32_t) * (size_t)np) != hipSuccess ||
        hipMalloc((void **)&t.bstart, sizeof(int32_t) * ((size_t)c->g.n + 1)) != hipSuccess)
      throw std::bad_alloc();
  }
  const int seg_len = std::min(t.every, t.nsteps);
  // the start of every step of the segment and the result of its last; when
  // memory allows, also the four stage states of every step of the segment,
  // kept from the segment's one recomputation (else each step's stages are
  // recomputed from its start when it is reversed)
  if (!t.seg && hipMalloc((void **)&t.seg, sizeof(double) * nq * (size_t)np * (seg_len + 1)) != hipSuccess)
    throw std::bad_alloc();
  if (!t.segstage && !t.segstage_tried) {
    t.segstage_tried = true;
    if (hipMalloc((void **)&t.segstage, sizeof(double) * 4 * nq * (size_t)np * seg_len) != hipSuccess) {
      (void)hipGetLastError();
      t.segstage = nullptr;
    }
  }
  const double *in[7] = {};
  double *out[7] = {};
  auto start_of = [&](int j, int q) { return t.seg + ((long)j * nq + q) * np; };
  auto stages_of = [&](int j) { return t.segstage + (long)j * 4 * nq * np; };
  int seg_k = -1;
  for (int s = t.nsteps - 1; s >= 0; --s) {
    const int k0 = s / t.every, base = k0 * t.every;
    if (k0 != seg_k) {
      if (gas && gas->segment) gas->segment(k0);  // the segment's gas tape, replayed from its checkpoint
      for (int q = 0; q < nq; ++q) launch_copy(c, start_of(0, q), t.ckpt + ((long)k0 * nq + q) * np, np);
      const int last = t.segstage ? s : s - 1;
      for (int r = base; r <= last; ++r) {
        for (int q = 0; q < nq; ++q) {
          in[q] = start_of(r - base, q);
          out[q] = start_of(r - base + 1, q);
        }
        tpx_launch<k_adj_step_fwd>(grid1(np), kBlock, 0, c->stream,
                                   step_args(c, r, in, out, t.segstage ? stages_of(r - base) : nullptr));
        TPX_LAUNCH_CHECK();
      }
      seg_k = k0;
    }
    // the stage states of step s and its result
    double *stg = t.stage, *yres = t.y;
    if (t.segstage) {
      stg = stages_of(s - base);
      yres = start_of(s - base + 1, 0);
    } else {
      for (int q = 0; q < nq; ++q) in[q] = start_of(s - base, q);
      tpx_launch<k_adj_step_fwd>(grid1(np), kBlock, 0, c->stream, step_args(c, s, in, y, t.stage));
      TPX_LAUNCH_CHECK();
    }
    if (!gas) {
      AdjArgs a{};
      a.g = c->g;
      a.gas = t.gas_of(s, c->g.n);
      a.stage = stg;
      for (int q = 0; q < 6; ++q) a.lam[q] = lam[q];
      a.dtau = t.dtau;
      a.tau = pt.has_tau ? t.tau0 : nullptr;
      a.tp = c->Tp;
      for (int r = 0; r < 4; ++r) {
        a.c_pre[r] = c->p.dt * c->pre[r];
        a.c_post[r] = c->p.dt * c->post[r];
      }
      a.np = np;
      a.ncell = c->g.n;
      tpx_launch<k_adj_step_rev>(grid1(np), kBlock, 0, c->stream, a);
      TPX_LAUNCH_CHECK();
      continue;
    }
    // the particles' stages interleaved with the gas substeps, backwards
    gas->begin(s);
    for (int rk = 3; rk >= 0; --rk) {
      gas->part_a(s, rk);
      if (t.heat) {  // the exchange at the end of this substep, at the next stage's state
        const double *Y = rk < 3 ? stg + (long)(rk + 1) * nq * np : yres;
        heat_reverse(c, Y, gas->heat_mid(s, rk), t.p0t + (long)s * 4 + rk + 1, rk < 3 ? t.lnext : t.lam,
                     SC_ADJ_P0NEW);
      }
      const double *const *lrhs = gas->part_rk(s, rk);
      StageArgs a{};
      a.g = c->g;
      a.gas = t.gas_of(s, c->g.n) + (long)rk * 4 * c->g.n;
      a.stage = stg + (long)rk * nq * np;
      a.lend = t.lam;
      a.lb = t.lb;
      a.lnext = t.lnext;
      a.dtau = t.dtau;
      a.lug = t.lug;
      for (int d = 0; d < 3; ++d) a.ls[d] = gas->a_drag != 0.0 ? lrhs[d] : nullptr;
      a.a_drag = gas->a_drag;
      a.tau = pt.has_tau ? t.tau0 : nullptr;
      a.tp = c->Tp;
      a.c_pre = c->p.dt * c->pre[rk];
      a.c_post = c->p.dt * c->post[rk];
      a.last = rk == 3;
      a.nq = nq;
      if (t.heat) {
        a.lh = t.lh;
        a.deps = t.deps;
        a.alpha_eps = 0.25 * TPX_PI * c->p.dp * c->p.dp * t.I0[(size_t)s];
        a.iS = 1. / (c->mp * c->p.cvp);
      }
      a.np = np;
      a.ncell = c->g.n;
      tpx_launch<k_adj_stage_rev>(grid1(np), kBlock, 0, c->stream, a);
      TPX_LAUNCH_CHECK();
      // the gather's transpose of lug onto the faces, pulled per cell
      bin_positions(c, a.stage, false);
      tpx_launch<k_adj_gather_pull>(grid1d(c->g.n, kBlock), kBlock, 0, c->stream, c->g, a.stage,
                                    (const double *)t.lug, (const int32_t *)t.bstart, (const int32_t *)t.bperm,
                                    t.lup, np);
      TPX_LAUNCH_CHECK();
      gas->part_b(s, rk, t.lup);
    }
    tpx_launch<k_adj_stage_end>(grid1(nq * np), kBlock, 0, c->stream, (const double *)t.lb, (const double *)t.lnext,
                                t.lam, nq * np);
    TPX_LAUNCH_CHECK();
  }
  // the exchange that made the recording's first RHS_T (at its initial state)
  if (t.heat) heat_reverse(c, t.ckpt, gas->heat_init, t.p0t, t.lam, SC_ADJ_P0I);
}

}  // namespace tpx
