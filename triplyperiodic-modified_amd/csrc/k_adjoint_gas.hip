// k_adjoint_gas.hip -- hand-written adjoint of the gas substep (gfx950).
//
// The reverse sweep of one RK4 substep of the gas (grid.cpp:608-614,
// 776-863), from the substep as the recording kept it or as recomputed from
// its step's start.  With
//   F  = I_C2F(Rho_int),  U = RU_int / F                       (G3)
//   RHS = a_drag S + mu/3 grad div U + mu del2 U - conv(U, RU_int)
//         + A RU_int + g (F - rho0)                            (G4)
// (S the particles' drag deposit under two-way coupling: its adjoint
// a_drag lam_RHS goes back to the particles' stage, k_adjoint.hip)
//   D_rho = div RU_int                                         (G1)
//   F' = I_C2F(Rho_new),  C = 1/F',  cc = -1/I_F2C(C)
//   div   = k/cp del2(1/Rho_new) + c0 Q',  c0 = -R/(cp P0_new Vcell)  (G5)
//   RHS_P = Z((div(RU_new/F') - div) / (dt pre))               (G6)
//   P_{i+1} = P_i + M(cc (div(C grad P_i) - RHS_P)),  i < n     (G7)
//   RU_np1 -= dt post grad P,  RU_new -= dt pre grad P          (G10)
// (Z the mean-zero projection, M the spectral solve of poisson.cpp:413-462,
// which is real and symmetric), the transposes used here are
//   div^T = -grad, I_C2F^T = I_F2C along the same axis, del2 and grad div
//   symmetric, M^T = M, div(C grad .) symmetric for fixed C,
// and the Poisson iterations are reversed with their recorded count:
//   nu = M lam_P,  mu = cc nu,  lam_P += div(C grad mu),  lam_RHS_P -= mu,
//   lam_C -= grad mu . grad P_i,  lam_cc += nu (div(C grad P_i) - RHS_P).
// With the heat exchange (Q = RHS_Part_Temp, Q' = Q - [cooling] mean Q) the
// divergence's Q term gives lam_Q' = -c0 nu (nu the adjoint of -div) and
// lam_P0_new += (c0/P0_new) sum nu Q', and Update_P0 (grid.cpp:816-830),
//   dP0 = -(R/cv) (sum Q / ncell) / Vcell,  P0_np1 += dt post dP0,
//   P0_new = P0 + dt pre dP0 (P0_np1 at the fourth substep),
// is reversed on the device scalars; the particles' side of the exchange
// (the deposit of RHS_T and the gather of T = P0/(R Rho)) is in k_adjoint.hip.
// Face fields follow the forward layout: component a at index x is the face
// between cells x - e_a and x.  Every kernel is one thread per cell (and the
// three faces at its index) over the periodic box.
#include <math.h>

#include <stdexcept>

#include "tpx_internal.h"

namespace tpx {

namespace {

constexpr int kBlock = 256;

// a cell and its periodic neighbours
struct Cell {
  int i, j, k, mx, my, mz, nx;
  long nxy;
  __device__ __forceinline__ Cell(const Geo &g, long c) {
    mx = g.nx - 1; my = g.ny - 1; mz = g.nz - 1; nx = g.nx; nxy = g.nxy;
    const int lx = __builtin_ctz(g.nx), ly = __builtin_ctz(g.ny);
    i = (int)(c & mx);
    j = (int)((c >> lx) & my);
    k = (int)(c >> (lx + ly));
  }
  __device__ __forceinline__ long off(int di, int dj, int dk) const {
    return (long)((k + dk) & mz) * nxy + (long)((j + dj) & my) * nx + ((i + di) & mx);
  }
  // shifted by da along axis a and db along axis b
  __device__ __forceinline__ long at(int a, int da, int b = 0, int db = 0) const {
    return off((a == 0) * da + (b == 0) * db, (a == 1) * da + (b == 1) * db, (a == 2) * da + (b == 2) * db);
  }
};

__device__ __forceinline__ double inv_d(const Geo &g, int a) { return a == 0 ? g.idx : (a == 1 ? g.idy : g.idz); }
__device__ __forceinline__ double inv_d2(const Geo &g, int a) { return a == 0 ? g.idx2 : (a == 1 ? g.idy2 : g.idz2); }

// 1/x as the forward Poisson iteration forms it (reciprocal estimate and two
// Newton steps): C and cc here are the iteration's own, so that the spectral
// solve below sees the residual it is given
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// block sum of v into partials[blockIdx.x]
__device__ __forceinline__ void block_sum(double v, double *partials) {
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// block sums of u and v into pu[blockIdx.x] and pv[blockIdx.x]
__device__ __forceinline__ void block_sum2(double u, double v, double *pu, double *pv) {
  __shared__ double sh[2][kBlock];
  sh[0][threadIdx.x] = u;
  sh[1][threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pu[blockIdx.x] = sh[0][0];
    pv[blockIdx.x] = sh[1][0];
  }
}

// ---- projection (G10) ------------------------------------------------------
struct GradPArgs {
  Geo g;
  const double *np1[3], *nw[3];
  const double *rho;      // Rho_new
  double *lp;
  double *cc;             // out: cc of every cell, for the Poisson kernels
  double c_np1, c_new;    // dt post, and dt pre (dt post at the last substep)
};

// lam_P += div(dt post lam_np1 + dt pre lam_new) (grad^T = -div), and the
// substep's cc = -1/I_F2C(C), C = 1/I_C2F(Rho_new)
struct k_adj_grad_p {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const GradPArgs &a) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (c >= a.g.n) return;
    const Cell x(a.g, c);
    double s = 0.0, csum = 0.0;
    const double rc = a.rho[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const long p = x.at(d, 1), m = x.at(d, -1);
      const double q0 = a.c_np1 * a.np1[d][c] + a.c_new * a.nw[d][c];
      const double q1 = a.c_np1 * a.np1[d][p] + a.c_new * a.nw[d][p];
      s += (q1 - q0) * inv_d(a.g, d);
      const double c0 = rcp_nr((rc + a.rho[m]) * 0.5), c1 = rcp_nr((a.rho[p] + rc) * 0.5);
      csum += (c0 + c1) * 0.5;
    }
    a.cc[c] = -rcp_nr((1. / 3.) * csum);
    a.lp[c] = a.lp[c] + s;
  }
};

// ---- Poisson solve (G7) ----------------------------------------------------
struct PoisArgs {
  Geo g;
  const double *rho;       // Rho_new
  const double *cc;
  const double *nu;        // M lam_P, the spectral solve of the iteration being reversed
  const double *Pi;        // P before that iteration
  const double *rhs;       // raw RHS_Pois
  const double *rhs_sum;   // its sum (one value)
  double *lp;              // lam_P
  double *lrhs, *lcc;
  double *lrho;            // lam_Rho_new
  double *lF[3];           // lam_F' (pois_c's first write)
  double *partials;        // block sums of lam_RHS after the last reversed iteration
  int first;               // the first reversed iteration writes the accumulators
  int last;                // the last one (the solve's first iteration)
};

// One Poisson iteration reversed, after its spectral solve nu = M lam_P:
//   mu = cc nu,  lam_RHS -= mu,  lam_cc += nu (div(C grad P_i) - RHS_P),
//   lam_P += div(C grad mu),
// and the coefficient's share lam_C = -grad mu . grad P_i on each face taken
// at once through C = 1/F' (lam_F' = -lam_C C^2) and F' = I_C2F(Rho_new)
// (each face to its two cells, half each) into lam_Rho_new: the cell gets
// 1/2 (C^2 grad mu grad P) of its six faces.  C is formed from Rho_new as the
// forward iteration forms it; mu at the neighbours from their cc and nu.
struct k_adj_pois_it {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const PoisArgs &a) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    double lr = 0.0;
    if (c < a.g.n) {
      const Cell x(a.g, c);
      const double nuc = a.nu[c];
      const double muc = a.cc[c] * nuc;
      const double rc = a.rho[c], pc = a.Pi[c];
      double dmu = 0.0, dpi = 0.0, lrho = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const long m = x.at(d, -1), p = x.at(d, 1);
        const double id = inv_d(a.g, d);
        const double c0 = rcp_nr((rc + a.rho[m]) * 0.5), c1 = rcp_nr((a.rho[p] + rc) * 0.5);
        const double mum = a.cc[m] * a.nu[m], mup = a.cc[p] * a.nu[p];
        const double gm0 = (muc - mum) * id, gm1 = (mup - muc) * id;
        const double gp0 = (pc - a.Pi[m]) * id, gp1 = (a.Pi[p] - pc) * id;
        dmu += (c1 * gm1 - c0 * gm0) * id;
        dpi += (c1 * gp1 - c0 * gp0) * id;
        lrho += 0.5 * ((c0 * c0) * (gm0 * gp0) + (c1 * c1) * (gm1 * gp1));
      }
      const double mean = a.rhs_sum[0] / (double)a.g.n;
      const double rr = dpi - (a.rhs[c] - mean);
      lr = (a.first ? 0.0 : a.lrhs[c]) - muc;
      a.lrhs[c] = lr;
      a.lcc[c] = (a.first ? 0.0 : a.lcc[c]) + nuc * rr;
      a.lp[c] = a.lp[c] + dmu;
      a.lrho[c] = a.lrho[c] + lrho;
    }
    if (a.last) block_sum(lr, a.partials);
  }
};

// cc's dependence on C, then C = 1/F': lam_F' = -lam_C C^2 with lam_C the
// cc share (first write of lam_F')
struct k_adj_pois_c {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const PoisArgs &a) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (c >= a.g.n) return;
    const Cell x(a.g, c);
    const double cc0 = a.cc[c], rc = a.rho[c];
    const double w0 = cc0 * cc0 * a.lcc[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const long m = x.at(d, -1);
      const double ccm = a.cc[m];
      const double lc = (w0 + ccm * ccm * a.lcc[m]) * (1. / 6.);
      const double C = rcp_nr((rc + a.rho[m]) * 0.5);
      a.lF[d][c] = -lc * C * C;
    }
  }
};

// ---- Poisson right-hand side and conduction term (G5, G6) -------------------
struct RhsPArgs {
  Geo g;
  const double *rho;      // Rho_new
  const double *ru_new[3];
  const double *lrhs;
  const double *scal;     // SC_ADJ_SUM: sum of lam_RHS
  double s, kappa;        // 1/(dt pre), k/cp
  double *lnew[3], *lF[3], *lrho;
  // heat exchange (Q null: none)
  const double *Q, *q_sum, *p0n;
  double *lq;
  double *pa, *pb;        // per block: lam_P0_new and lam_(sum Q) terms
  double R, cp, Vcell;
  int cooling;
};

struct k_adj_rhs_pois {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const RhsPArgs &a) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    double ta = 0.0, tb = 0.0;
    if (c < a.g.n) {
      const Cell x(a.g, c);
      const double mean = a.scal[SC_ADJ_SUM] / (double)a.g.n;
      auto nu = [&](long p) { return a.s * (a.lrhs[p] - mean); };
      const double nc = nu(c), rc = a.rho[c];
      double lap = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const long m = x.at(d, -1), p = x.at(d, 1);
        const double nm = nu(m);
        lap += (nu(p) - 2 * nc + nm) * inv_d2(a.g, d);
        const double F = (rc + a.rho[m]) * 0.5;
        const double lu = -((nc - nm) * inv_d(a.g, d));
        a.lnew[d][c] += lu / F;
        a.lF[d][c] += -lu * a.ru_new[d][c] / (F * F);
      }
      a.lrho[c] += a.kappa * lap / (rc * rc);
      if (a.Q) {
        const double P0n = a.p0n[0];
        const double c0 = -a.R / (a.cp * P0n * a.Vcell);
        const double qp = a.Q[c] - (a.cooling ? a.q_sum[0] / (double)a.g.n : 0.0);
        const double lqp = -c0 * nc;
        a.lq[c] = lqp;
        ta = (c0 / P0n) * nc * qp;
        tb = a.cooling ? -lqp / (double)a.g.n : 0.0;
      }
    }
    if (a.Q) block_sum2(ta, tb, a.pa, a.pb);
  }
};

// lam_rho += I_F2C-transpose of a face adjoint: each face feeds its two cells
__device__ __forceinline__ double fold_at(const Cell &x, long c, const double *const lF[3]) {
  return 0.5 * ((lF[0][c] + lF[0][x.at(0, 1)]) + (lF[1][c] + lF[1][x.at(1, 1)]) + (lF[2][c] + lF[2][x.at(2, 1)]));
}

struct k_adj_fold {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const double *lF0, const double *lF1, const double *lF2,
                                             double *out) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (c >= g.n) return;
    const Cell x(g, c);
    const double *lF[3] = {lF0, lF1, lF2};
    out[c] = fold_at(x, c, lF);
  }
};

// out += the fold of a face adjoint
struct k_adj_fold_add {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const double *lF0, const double *lF1, const double *lF2,
                                             double *out) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (c >= g.n) return;
    const Cell x(g, c);
    const double *lF[3] = {lF0, lF1, lF2};
    out[c] = out[c] + fold_at(x, c, lF);
  }
};

// ---- RK4 bookkeeping of the momentum and density updates (G4, G1) ---------
struct RkArgs {
  Geo g;
  double *lnp1[3], *lnew[3], *lru[3], *lrhs[3];
  double *lr, *lrnp1, *lrnew, *ld;
  const double *lF[3];       // the Poisson and right-hand side's face adjoint, folded into lam_Rho_new first
  const double *rui[3], *rho_int;
  double *U[3];              // out: U = RU_int / I_C2F(Rho_int) for the momentum kernels
  double c_post, c_pre;
  int last;  // the fourth substep: new = np1
};

struct k_adj_rk {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const RkArgs &a) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (c >= a.g.n) return;
    const Cell x(a.g, c);
    const double rc = a.rho_int[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double ln = a.lnew[d][c];
      if (a.last) {
        const double l1 = a.lnp1[d][c] + ln;
        a.lnp1[d][c] = l1;
        a.lrhs[d][c] = a.c_post * l1;
      } else {
        a.lrhs[d][c] = a.c_post * a.lnp1[d][c] + a.c_pre * ln;
        a.lru[d][c] += ln;
      }
      a.U[d][c] = a.rui[d][c] / ((rc + a.rho_int[x.at(d, -1)]) * 0.5);
    }
    const double lrn = a.lrnew[c] + fold_at(x, c, a.lF);
    if (a.last) {
      const double l1 = a.lrnp1[c] + lrn;
      a.lrnp1[c] = l1;
      a.ld[c] = -a.c_post * l1;
    } else {
      a.ld[c] = -a.c_post * a.lrnp1[c] - a.c_pre * lrn;
      a.lr[c] += lrn;
    }
  }
};

// ---- momentum right-hand side (G4) -----------------------------------------
struct MomArgs {
  Geo g;
  const double *lrhs[3];
  const double *rui[3];
  const double *rho;        // Rho_int
  const double *U[3];       // RU_int / I_C2F(Rho_int)
  const double *lup[3];     // the particles' gather adjoint on the faces, or null
  const double *ld;         // adjoint of div RU_int (G1)
  double *lu[3];            // adjoint of U
  double *lint[3], *lF[3];  // adjoint of RU_int, of F (first writes)
  double *partials;         // per block: sum lam_RHS . RU_int (the forcing coefficient's adjoint)
  double b_visc, c_visc, A, gv[3];
};

// adjoint of U = RU_int/F: viscous terms, the convective term's U factor
// and the particles' gather
struct k_adj_mom_u {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const MomArgs &a) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (c >= a.g.n) return;
    const Cell x(a.g, c);
    const Geo &g = a.g;
    // div lam_RHS at the cell and at its lower neighbour along each axis
    auto divl = [&](int di, int dj, int dk) {
      const long q = x.off(di, dj, dk);
      return (a.lrhs[0][x.off(di + 1, dj, dk)] - a.lrhs[0][q]) * g.idx +
             (a.lrhs[1][x.off(di, dj + 1, dk)] - a.lrhs[1][q]) * g.idy +
             (a.lrhs[2][x.off(di, dj, dk + 1)] - a.lrhs[2][q]) * g.idz;
    };
    const double dv0 = divl(0, 0, 0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {  // component d of U
      const double *L = a.lrhs[d];
      const double lc = L[c];
      double v = a.b_visc * (dv0 - divl(-(d == 0), -(d == 1), -(d == 2))) * inv_d(g, d);
      double lap = 0.0;
#pragma unroll
      for (int b = 0; b < 3; ++b) lap += (L[x.at(b, 1)] - 2 * lc + L[x.at(b, -1)]) * inv_d2(g, b);
      v += a.c_visc * lap;
      // -div_F2C(I_C2F(U_d) * I_d(RU_int)): the U_d factor, through every axis b
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double idb = inv_d(g, b);
        const double *R = a.rui[b];
        const long pb = x.at(b, 1);
        const double w0 = (lc - L[x.at(b, -1)]) * idb;
        const double w1 = (L[pb] - lc) * idb;
        const double d0 = (R[c] + R[x.at(d, -1)]) * 0.5;
        const double d1 = (R[pb] + R[x.at(b, 1, d, -1)]) * 0.5;
        v += 0.5 * (w0 * d0 + w1 * d1);
      }
      if (a.lup[0]) v += a.lup[d][c];
      a.lu[d][c] = v;
    }
  }
};

// adjoint of RU_int (forcing, convection's RU factor, U = RU/F, div RU_int of
// the density update) and of F (body force, U = RU/F); the forcing
// coefficient's adjoint as block partial sums
struct k_adj_mom_r {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const MomArgs &a) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    double part = 0.0;
    if (c < a.g.n) {
      const Cell x(a.g, c);
      const Geo &g = a.g;
#pragma unroll
      for (int b = 0; b < 3; ++b) {  // component b of RU_int, face b at c
        const double idb = inv_d(g, b);
        const long mb = x.at(b, -1);
        const double iF = 1.0 / ((a.rho[c] + a.rho[mb]) * 0.5);
        const double lrb = a.lrhs[b][c], lub = a.lu[b][c];
        double v = lub * iF + a.A * lrb - (a.ld[c] - a.ld[mb]) * idb;
        // the RU_b factor of every component d's convective flux along b
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double *L = a.lrhs[d], *Ud = a.U[d];
          const long pd = x.at(d, 1), pdmb = x.at(d, 1, b, -1);
          const double w0 = (L[c] - L[mb]) * idb;
          const double w1 = (L[pd] - L[pdmb]) * idb;
          // U_d on its faces at c, c - e_b, c + e_d, c + e_d - e_b
          v += 0.5 * (w0 * (Ud[c] + Ud[mb]) * 0.5 + w1 * (Ud[pd] + Ud[pdmb]) * 0.5);
        }
        a.lint[b][c] = v;
        a.lF[b][c] = a.gv[b] * lrb - lub * a.U[b][c] * iF;
        part += lrb * a.rui[b][c];
      }
    }
    block_sum(part, a.partials);
  }
};

// acc[0] += sum of partials (one block)
struct k_adj_acc {
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

// ---- step boundary (TimeAdvance, make_mean_U0) -----------------------------
// RU = RU_np1 and Rho = Rho_np1 at the start of the next step: their adjoints
// join the np1 accumulators (and the base accumulators restart at zero)
struct k_adj_boundary {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(long n, double *l0, double *l1, double *l2, double *b0, double *b1,
                                             double *b2, double *lr, double *br) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (c >= n) return;
    l0[c] += b0[c]; b0[c] = 0.0;
    l1[c] += b1[c]; b1[c] = 0.0;
    l2[c] += b2[c]; b2[c] = 0.0;
    lr[c] += br[c]; br[c] = 0.0;
  }
};

// f -= scal[slot] / n: the transpose of make_mean_U0's mean removal
struct k_adj_sub_mean {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(long n, double *f, const double *scal, int slot) {
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (c >= n) return;
    f[c] -= scal[slot] / (double)n;
  }
};

// Update_P0 reversed for one substep (step false), or the step boundary
// (step true: P0 = P0_np1 at the next step's start, P0_np1 = P0 at this
// one's): the adjoints of P0, P0_np1 and this substep's P0_new, and that of
// the sum of RHS_Part_Temp the update read (written, not accumulated: the
// heat exchange that produced it consumes it next)
struct k_adj_p0 {
  static constexpr int kBounds = 64;
  static __device__ __forceinline__ void run(double *scal, bool step, int last, int cooling, double c_np1,
                                             double c_new, double dsum) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (step) {
      scal[SC_ADJ_P0NP1] += scal[SC_ADJ_P0];
      scal[SC_ADJ_P0] = 0.0;
      return;
    }
    const double ln = scal[SC_ADJ_P0NEW];
    double ld;
    if (last) {
      const double l1 = scal[SC_ADJ_P0NP1] + ln;
      scal[SC_ADJ_P0NP1] = l1;
      ld = c_np1 * l1;
    } else {
      ld = c_np1 * scal[SC_ADJ_P0NP1] + c_new * ln;
      scal[SC_ADJ_P0] += ln;
    }
    scal[SC_ADJ_LSQ] = cooling ? 0.0 : ld * dsum;
    scal[SC_ADJ_P0NEW] = 0.0;
  }
};

unsigned cells(const tpx_ctx *c) { return grid1d(c->g.n, kBlock); }

}  // namespace

void adjoint_gas_boundary(tpx_ctx *c, GasAdjoint &ga, bool heat) {
  const long n = c->g.n;
  if (heat) {
    tpx_launch<k_adj_p0>(1, 64, 0, c->stream, c->scal, true, 0, 0, 0.0, 0.0, 0.0);
    TPX_LAUNCH_CHECK();
  }
  tpx_launch<k_adj_boundary>(cells(c), kBlock, 0, c->stream, n, ga.lnp1[0], ga.lnp1[1], ga.lnp1[2], ga.lru[0],
                             ga.lru[1], ga.lru[2], ga.lrnp1, ga.lr);
  TPX_LAUNCH_CHECK();
  for (int d = 0; d < 3; ++d) {
    launch_reduce(c, ga.lnp1[d], 0, SC_ADJ_M0 + d);
    tpx_launch<k_adj_sub_mean>(cells(c), kBlock, 0, c->stream, n, ga.lnp1[d], (const double *)c->scal, SC_ADJ_M0 + d);
    TPX_LAUNCH_CHECK();
  }
}

void adjoint_gas_substep_a(tpx_ctx *c, GasAdjoint &ga, const GasSubstep &s, const Coef &k) {
  const unsigned nb = cells(c);
  if (s.iters < 1) throw std::logic_error("adjoint: a Poisson solve without iterations");
  // G11: the int state after this substep is its new state
  for (int d = 0; d < 3; ++d) std::swap(ga.lnew[d], ga.lint[d]);
  std::swap(ga.lrnew, ga.lrint);
  const bool last = k.rk == 3;
  // G10 (and the Poisson coefficient, and the first spectral solve's input)
  {
    GradPArgs a{};
    a.g = c->g;
    for (int d = 0; d < 3; ++d) { a.np1[d] = ga.lnp1[d]; a.nw[d] = ga.lnew[d]; }
    a.rho = s.rho_new;
    a.lp = ga.lp;
    a.cc = ga.cc;
    a.c_np1 = k.dt * k.post;
    a.c_new = last ? k.dt * k.post : k.dt * k.pre;
    tpx_launch<k_adj_grad_p>(nb, kBlock, 0, c->stream, a);
    TPX_LAUNCH_CHECK();
  }
  // G7, iterations in reverse: each applies the spectral step to lam_P
  // (M = M^T) and then the rest of the iteration's transpose
  PoisArgs pa{};
  pa.g = c->g;
  pa.rho = s.rho_new;
  for (int d = 0; d < 3; ++d) pa.lF[d] = ga.lF[d];
  pa.cc = ga.cc;
  pa.nu = ga.nu;
  pa.rhs = s.rhs;
  pa.rhs_sum = s.rhs_sum;
  pa.lp = ga.lp;
  pa.lrhs = ga.lrhs_p;
  pa.lcc = ga.lcc;
  pa.lrho = ga.lrnew;
  pa.partials = c->partials;
  for (int i = s.iters - 1; i >= 0; --i) {
    pa.Pi = s.P[i];
    pa.first = i == s.iters - 1;
    pa.last = i == 0;
    launch_spectral_apply(c, ga.lp, ga.nu, ga.scal);
    tpx_launch<k_adj_pois_it>(nb, kBlock, 0, c->stream, pa);
    TPX_LAUNCH_CHECK();
  }
  launch_finish_sum(c, (int)nb, 0, SC_ADJ_SUM);  // sum of lam_RHS_Pois from the last k_adj_pois_it
  tpx_launch<k_adj_pois_c>(nb, kBlock, 0, c->stream, pa);
  TPX_LAUNCH_CHECK();
  // G6 and G5
  {
    RhsPArgs a{};
    a.g = c->g;
    a.rho = s.rho_new;
    for (int d = 0; d < 3; ++d) { a.ru_new[d] = s.ru_new[d]; a.lnew[d] = ga.lnew[d]; a.lF[d] = ga.lF[d]; }
    a.lrhs = ga.lrhs_p;
    a.scal = c->scal;
    a.s = 1. / (k.dt * k.pre);
    a.kappa = k.k_over_cp;
    a.lrho = ga.lrnew;
    a.Q = s.Q;
    if (s.Q) {
      a.q_sum = s.q_sum;
      a.p0n = s.p0n;
      a.lq = ga.lq;
      a.pa = c->partials;
      a.pb = c->partials + c->n_partials;
      a.R = k.R;
      a.cp = k.cp;
      a.Vcell = k.Vcell;
      a.cooling = k.cooling;
    }
    tpx_launch<k_adj_rhs_pois>(nb, kBlock, 0, c->stream, a);
    TPX_LAUNCH_CHECK();
    if (s.Q) {
      tpx_launch<k_adj_acc>(1, 1024, 0, c->stream, (const double *)a.pa, (long)nb, c->scal + SC_ADJ_P0NEW);
      TPX_LAUNCH_CHECK();
      tpx_launch<k_adj_acc>(1, 1024, 0, c->stream, (const double *)a.pb, (long)nb, c->scal + SC_ADJ_LSQ);
      TPX_LAUNCH_CHECK();
    }
  }
}

const double *const *adjoint_gas_substep_rk(tpx_ctx *c, GasAdjoint &ga, const GasSubstep &s, const Coef &k) {
  const unsigned nb = cells(c);
  const bool last = k.rk == 3;
  // G4 and G1: lam_Rho_new complete (fold of the face adjoint, and the heat
  // exchange's gather of T), the RK4 updates, then the right-hand sides
  {
    RkArgs a{};
    a.g = c->g;
    for (int d = 0; d < 3; ++d) {
      a.lnp1[d] = ga.lnp1[d]; a.lnew[d] = ga.lnew[d]; a.lru[d] = ga.lru[d]; a.lrhs[d] = ga.lrhs[d];
      a.lF[d] = ga.lF[d]; a.rui[d] = s.ru_int[d]; a.U[d] = ga.U[d];
    }
    a.lr = ga.lr; a.lrnp1 = ga.lrnp1; a.lrnew = ga.lrnew; a.ld = ga.ld;
    a.rho_int = s.rho_int;
    a.c_post = k.dt * k.post;
    a.c_pre = k.dt * k.pre;
    a.last = last;
    tpx_launch<k_adj_rk>(nb, kBlock, 0, c->stream, a);
    TPX_LAUNCH_CHECK();
  }
  return ga.lrhs;
}

void adjoint_gas_substep_b(tpx_ctx *c, GasAdjoint &ga, const GasSubstep &s, const Coef &k) {
  const long n = c->g.n;
  const unsigned nb = cells(c);
  MomArgs m{};
  m.g = c->g;
  for (int d = 0; d < 3; ++d) {
    m.lrhs[d] = ga.lrhs[d];
    m.rui[d] = s.ru_int[d];
    m.U[d] = ga.U[d];
    m.lup[d] = s.lup ? s.lup + (long)d * n : nullptr;
    m.lu[d] = ga.lu[d];
    m.lint[d] = ga.lint[d];
    m.lF[d] = ga.lF[d];
    m.gv[d] = k.g[d];
  }
  m.rho = s.rho_int;
  m.ld = ga.ld;
  m.partials = c->partials;
  m.b_visc = k.b_visc;
  m.c_visc = k.c_visc;
  m.A = k.A;
  tpx_launch<k_adj_mom_u>(nb, kBlock, 0, c->stream, m);
  TPX_LAUNCH_CHECK();
  tpx_launch<k_adj_mom_r>(nb, kBlock, 0, c->stream, m);
  TPX_LAUNCH_CHECK();
  tpx_launch<k_adj_acc>(1, 1024, 0, c->stream, (const double *)c->partials, (long)nb, c->scal + SC_ADJ_A);
  TPX_LAUNCH_CHECK();
  tpx_launch<k_adj_fold>(nb, kBlock, 0, c->stream, c->g, (const double *)ga.lF[0], (const double *)ga.lF[1],
                         (const double *)ga.lF[2], ga.lrint);
  TPX_LAUNCH_CHECK();
  if (s.Q) {  // G2
    tpx_launch<k_adj_p0>(1, 64, 0, c->stream, c->scal, false, k.rk == 3 ? 1 : 0, k.cooling, k.dt * k.post,
                         k.dt * k.pre, -k.R / k.cv / k.Vcell / (double)n);
    TPX_LAUNCH_CHECK();
  }
}

void adjoint_poisson_check(tpx_ctx *c, int op, const double *rho, const double *x, const double *y, double *ax,
                           double *aty, double *atr, double *w) {
  const long n = c->g.n;
  const unsigned nb = cells(c);
  double *cc = w, *nu = w + n, *lrhs = w + 2 * n, *lcc = w + 3 * n, *zero = w + 4 * n;
  double *lF[3] = {w + 5 * n, w + 6 * n, w + 7 * n};
  double *scal = w + 8 * n;
  launch_fill(c, w, 0.0, 8 * n + SC_COUNT);
  if (op == 0) {
    launch_spectral_apply(c, x, ax, scal);
    launch_spectral_apply(c, y, aty, scal);
    return;
  }
  // one forward iteration from P = x with a zero right-hand side
  launch_copy(c, ax, x, n);
  launch_poisson_iteration_on(c, rho, ax, zero, scal);
  launch_fill(c, scal, 0.0, SC_COUNT);  // its loop test, undone for the spectral step below
  // the same iteration reversed for lam = y: cc, nu = M y, then the kernels
  // of the sweep (the projection term is zero here)
  GradPArgs ga{};
  ga.g = c->g;
  for (int d = 0; d < 3; ++d) { ga.np1[d] = zero; ga.nw[d] = zero; }
  ga.rho = rho;
  ga.lp = aty;
  ga.cc = cc;
  launch_copy(c, aty, y, n);
  tpx_launch<k_adj_grad_p>(nb, kBlock, 0, c->stream, ga);
  TPX_LAUNCH_CHECK();
  launch_spectral_apply(c, aty, nu, scal);
  launch_fill(c, atr, 0.0, n);
  PoisArgs pa{};
  pa.g = c->g;
  pa.rho = rho;
  pa.cc = cc;
  pa.nu = nu;
  pa.Pi = x;
  pa.rhs = zero;
  pa.rhs_sum = zero;
  pa.lp = aty;
  pa.lrhs = lrhs;
  pa.lcc = lcc;
  pa.lrho = atr;
  for (int d = 0; d < 3; ++d) pa.lF[d] = lF[d];
  pa.partials = c->partials;
  pa.first = 1;
  pa.last = 1;
  tpx_launch<k_adj_pois_it>(nb, kBlock, 0, c->stream, pa);
  TPX_LAUNCH_CHECK();
  tpx_launch<k_adj_pois_c>(nb, kBlock, 0, c->stream, pa);
  TPX_LAUNCH_CHECK();
  tpx_launch<k_adj_fold_add>(nb, kBlock, 0, c->stream, c->g, (const double *)lF[0], (const double *)lF[1],
                             (const double *)lF[2], atr);
  TPX_LAUNCH_CHECK();
}

}  // namespace tpx
