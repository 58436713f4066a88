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

// The cell block of a work group: blocks renumbered so that each XCD takes
// one contiguous range of cells, and the stencils' neighbouring rows and
// planes are in that XCD's L2.  Block partial sums are stored under this
// number, so they cover the same cells at the same index as a plain
// numbering would.
__device__ __forceinline__ unsigned cell_block() { return xcd_block(blockIdx.x, gridDim.x); }

// block sum of v into partials[cell_block()]
__device__ __forceinline__ void block_sum(double v, double *partials) {
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[cell_block()] = sh[0];
}

// block sum of v into partials[cell_block()] for a group of any power-of-two
// width up to kBlock
__device__ __forceinline__ void block_sum_bd(double v, double *partials) {
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = (int)block_dim_x() / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[cell_block()] = sh[0];
}

// block sums of u and v into pu[cell_block()] and pv[cell_block()]
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
    pu[cell_block()] = sh[0][0];
    pv[cell_block()] = sh[1][0];
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
    const long c = (long)cell_block() * block_dim_x() + threadIdx.x;
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


// cc's dependence on C, then C = 1/F': lam_F' = -lam_C C^2 with lam_C the
// cc share (first write of lam_F')
struct k_adj_pois_c {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const PoisArgs &a) {
    const long c = (long)cell_block() * block_dim_x() + threadIdx.x;
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
  // non-null: the Poisson coefficient's share of lam_F' (k_adj_pois_c's) is
  // formed here from cc and lam_cc and lam_F' written, not accumulated
  const double *cc, *lcc;
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
    const long c = (long)cell_block() * block_dim_x() + threadIdx.x;
    double ta = 0.0, tb = 0.0;
    if (c < a.g.n) {
      const Cell x(a.g, c);
      const double mean = a.scal[SC_ADJ_SUM] / (double)a.g.n;
      auto nu = [&](long p) { return a.s * (a.lrhs[p] - mean); };
      const double nc = nu(c), rc = a.rho[c];
      const double w0 = a.lcc ? a.cc[c] * a.cc[c] * a.lcc[c] : 0.0;
      double lap = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const long m = x.at(d, -1), p = x.at(d, 1);
        const double nm = nu(m);
        lap += (nu(p) - 2 * nc + nm) * inv_d2(a.g, d);
        const double F = (rc + a.rho[m]) * 0.5;
        const double lu = -((nc - nm) * inv_d(a.g, d));
        a.lnew[d][c] += lu / F;
        if (a.lcc) {
          const double ccm = a.cc[m];
          const double lc = (w0 + ccm * ccm * a.lcc[m]) * (1. / 6.);
          const double C = rcp_nr((rc + a.rho[m]) * 0.5);
          const double pc = -lc * C * C;
          a.lF[d][c] = pc + -lu * a.ru_new[d][c] / (F * F);
        } else {
          a.lF[d][c] += -lu * a.ru_new[d][c] / (F * F);
        }
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
    const long c = (long)cell_block() * block_dim_x() + threadIdx.x;
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
    const long c = (long)cell_block() * block_dim_x() + threadIdx.x;
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
  const double *rho_int;
  double c_post, c_pre;
  int last;  // the fourth substep: new = np1
};

struct k_adj_rk {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const RkArgs &a) {
    const long c = (long)cell_block() * block_dim_x() + threadIdx.x;
    if (c >= a.g.n) return;
    const Cell x(a.g, c);
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
  const double *lup[3];     // the particles' gather adjoint on the faces, or null
  const double *ld;         // adjoint of div RU_int (G1)
  double *lint[3], *lF[3];  // adjoint of RU_int, of F (first writes)
  double *partials;         // per block: sum lam_RHS . RU_int (the forcing coefficient's adjoint)
  double b_visc, c_visc, A, gv[3];
};

// Tile of the momentum adjoint's z-march: TX x TY cells of a plane per work
// group (one thread each), KC planes marched per group.
struct AdjMomGeom { int TX, TY, KC; };

// The momentum right-hand side reversed, one cell per thread: the adjoint of
// component d of U = RU_int/F (viscous terms, the convective term's U factor
// through every axis b, the particles' gather), used at the cell only, then
// the adjoints of RU_int (forcing, convection's RU factor, U = RU/F, div
// RU_int of the density update) and of F (body force, U = RU/F), and the
// cell's share of the forcing coefficient's adjoint (lam_RHS . RU_int).
// val(q, ox, oy, oz) is field q at offset (ox, oy, oz) from the cell: q = 0-2
// lam_RHS, 3-5 RU_int, 6-8 U = RU_int/I_C2F(Rho_int), 9 Rho_int, 10 the
// adjoint of div RU_int.
template <class V>
__device__ __forceinline__ void adj_mom_cell(const MomArgs &a, const V &val, const double (&lup)[3], double (&lint)[3],
                                             double (&lF)[3], double &part) {
  const Geo &g = a.g;
  // div lam_RHS at offset (di, dj, dk)
  auto divl = [&](int di, int dj, int dk) {
    return (val(0, di + 1, dj, dk) - val(0, di, dj, dk)) * g.idx + (val(1, di, dj + 1, dk) - val(1, di, dj, dk)) * g.idy +
           (val(2, di, dj, dk + 1) - val(2, di, dj, dk)) * g.idz;
  };
  // field q at da along axis a and db along axis b
  auto at = [&](int q, int ax, int da, int bx = 0, int db = 0) {
    return val(q, (ax == 0) * da + (bx == 0) * db, (ax == 1) * da + (bx == 1) * db, (ax == 2) * da + (bx == 2) * db);
  };
  const double dv0 = divl(0, 0, 0);
  auto adj_u = [&](int d) {
    const double lc = val(d, 0, 0, 0);
    double v = a.b_visc * (dv0 - divl(-(d == 0), -(d == 1), -(d == 2))) * inv_d(g, d);
    double lap = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) lap += (at(d, b, 1) - 2 * lc + at(d, b, -1)) * inv_d2(g, b);
    v += a.c_visc * lap;
    // -div_F2C(I_C2F(U_d) * I_d(RU_int)): the U_d factor
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double idb = inv_d(g, b);
      const double w0 = (lc - at(d, b, -1)) * idb;
      const double w1 = (at(d, b, 1) - lc) * idb;
      const double d0 = (val(3 + b, 0, 0, 0) + at(3 + b, d, -1)) * 0.5;
      const double d1 = (at(3 + b, b, 1) + at(3 + b, b, 1, d, -1)) * 0.5;
      v += 0.5 * (w0 * d0 + w1 * d1);
    }
    if (a.lup[0]) v += lup[d];
    return v;
  };
#pragma unroll
  for (int b = 0; b < 3; ++b) {  // component b of RU_int, face b at the cell
    const double idb = inv_d(g, b);
    const double iF = 1.0 / ((val(9, 0, 0, 0) + at(9, b, -1)) * 0.5);
    const double lrb = val(b, 0, 0, 0), lub = adj_u(b);
    double v = lub * iF + a.A * lrb - (val(10, 0, 0, 0) - at(10, b, -1)) * idb;
    // the RU_b factor of every component d's convective flux along b
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double w0 = (val(d, 0, 0, 0) - at(d, b, -1)) * idb;
      const double w1 = (at(d, d, 1) - at(d, d, 1, b, -1)) * idb;
      // U_d on its faces at the cell, - e_b, + e_d, + e_d - e_b
      v += 0.5 * (w0 * (val(6 + d, 0, 0, 0) + at(6 + d, b, -1)) * 0.5 +
                  w1 * (at(6 + d, d, 1) + at(6 + d, d, 1, b, -1)) * 0.5);
    }
    lint[b] = v;
    lF[b] = a.gv[b] * lrb - lub * val(6 + b, 0, 0, 0) * iF;
    part += lrb * val(3 + b, 0, 0, 0);
  }
}

// The momentum adjoint as a z-march: a work group holds planes k and k + 1
// of the eleven fields on its tile (with a one-cell x/y halo) in LDS, U
// formed from RU_int and Rho_int as the tile is loaded; of plane k - 1 each
// thread keeps what the stencil reads there (its own column's values, and the
// x + 1 or y + 1 neighbour's for the fields read there) from when that plane
// was in LDS.  Every field value is read from HBM about once per march
// instead of up to seven times per cell from L2.
struct k_adj_mom {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const MomArgs &a, const AdjMomGeom zg) {
    constexpr int NF = 11;
    extern __shared__ double sm[];
    const Geo &g = a.g;
    const int TX = zg.TX, TY = zg.TY, KC = zg.KC;
    const int RW = TX + 2, RN = RW * (TY + 2);  // x from i0 - 1 to i0 + TX, y from j0 - 1 to j0 + TY
    const unsigned blk = cell_block();
    const int ntx = g.nx / TX, nty = g.ny / TY;
    const int i0 = (int)(blk % (unsigned)ntx) * TX, j0 = (int)((blk / (unsigned)ntx) % (unsigned)nty) * TY;
    const int k0 = (int)(blk / (unsigned)(ntx * nty)) * KC;
    const int tid = (int)threadIdx.x, bd = (int)block_dim_x(), tx = tid % TX, ty = tid / TX;
    const int mx = g.nx - 1, my = g.ny - 1, mz = g.nz - 1;
    const int i = i0 + tx, j = j0 + ty;
    const double *f[NF] = {a.lrhs[0], a.lrhs[1], a.lrhs[2], a.rui[0], a.rui[1], a.rui[2],
                           nullptr,   nullptr,   nullptr,   a.rho,    a.ld};
    auto gidx = [&](int x, int y, int z) { return (long)(z & mz) * g.nxy + (long)(y & my) * g.nx + (x & mx); };
    // field q at (x, y, z); U_d = RU_int_d / I_C2F(Rho_int) on face d
    // (q = 6 + d) is formed from RU_int and Rho_int as the forward forms it
    auto field = [&](int q, int x, int y, int z) -> double {
      if (q < 6 || q > 8) return f[q][gidx(x, y, z)];
      const int d = q - 6;
      const long o = gidx(x, y, z), m = gidx(x - (d == 0), y - (d == 1), z - (d == 2));
      return a.rui[d][o] / ((a.rho[o] + a.rho[m]) * 0.5);
    };
    // the tile point e of plane kz, its offset in the field
    auto T = [&](int slot, int q, int x, int y) -> double { return sm[(slot * NF + q) * RN + (y + 1) * RW + (x + 1)]; };
    auto load_tile = [&](int kz, int slot) {
      for (int e = tid; e < RN; e += bd) {
        const int x = i0 + e % RW - 1, y = j0 + e / RW - 1;
#pragma unroll
        for (int q = 0; q < NF; ++q) sm[(slot * NF + q) * RN + e] = field(q, x, y, kz);
      }
    };
    // plane k - 1: zc[q] at the cell, zx[0..2] of fields 0, 3, 6 at x + 1,
    // zy[0..2] of fields 1, 4, 7 at y + 1
    double zc[NF], zx[3], zy[3];
#pragma unroll
    for (int q = 0; q < NF; ++q) zc[q] = field(q, i, j, k0 - 1);
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      zx[u] = field(3 * u, i + 1, j, k0 - 1);
      zy[u] = field(3 * u + 1, i, j + 1, k0 - 1);
    }
    load_tile(k0, 0);
    load_tile(k0 + 1, 1);
    __syncthreads();
    double part = 0.0;
    int cur = 0;
    for (int kk = k0; kk < k0 + KC; ++kk) {
      const int nxt = cur ^ 1;
      const bool more = kk + 1 < k0 + KC;
      auto val = [&](int q, int ox, int oy, int oz) -> double {
        if (oz == 0) return T(cur, q, tx + ox, ty + oy);
        if (oz == 1) return T(nxt, q, tx + ox, ty + oy);
        if (ox == 0 && oy == 0) return zc[q];
        if (ox == 1 && oy == 0 && q < 9 && q % 3 == 0) return zx[q / 3];
        if (ox == 0 && oy == 1 && q < 9 && q % 3 == 1) return zy[q / 3];
        return __builtin_nan("");  // not read by the stencil
      };
      const long c = gidx(i, j, kk);
      double lup[3] = {0.0, 0.0, 0.0};
      if (a.lup[0]) {
#pragma unroll
        for (int d = 0; d < 3; ++d) lup[d] = a.lup[d][c];
      }
      double lint[3], lF[3];
      adj_mom_cell(a, val, lup, lint, lF, part);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        a.lint[d][c] = lint[d];
        a.lF[d][c] = lF[d];
      }
      // this plane becomes plane k - 1 of the next
#pragma unroll
      for (int q = 0; q < NF; ++q) zc[q] = T(cur, q, tx, ty);
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        zx[u] = T(cur, 3 * u, tx + 1, ty);
        zy[u] = T(cur, 3 * u + 1, tx, ty + 1);
      }
      __syncthreads();
      if (more) load_tile(kk + 2, cur);
      __syncthreads();
      cur = nxt;
    }
    block_sum_bd(part, a.partials);
  }
};

// One Poisson iteration reversed, after its spectral solve nu = M lam_P:
//   mu = cc nu,  lam_RHS -= mu,  lam_cc += nu (div(C grad P_i) - RHS_P),
//   lam_P += div(C grad mu),
// and the coefficient's share lam_C = -grad mu . grad P_i on each face taken
// at once through C = 1/F' (lam_F' = -lam_C C^2) and F' = I_C2F(Rho_new)
// (each face to its two cells, half each) into lam_Rho_new: the cell gets
// 1/2 (C^2 grad mu grad P) of its six faces.  C is formed from Rho_new as the
// forward iteration forms it; mu at the neighbours from their cc and nu.
// An LDS z-march: a group takes a TX x TY column of KC planes; each plane's
// nu, cc, Rho_new and P_i with a one-cell x/y halo come into LDS once (two
// slots, this plane and the next), the plane below's values at the cell stay
// in registers and the plane above's are read from the next slot: six loads a
// cell instead of 28 neighbour loads with one thread a cell (51 -> 44 us at
// 128^3).  The block sums of lam_RHS (the last reversed iteration) cover a
// column of cells each.
struct k_adj_pois_zm {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const PoisArgs &a, const AdjMomGeom zg) {
    constexpr int NF = 4;  // nu, cc, Rho_new, P_i
    extern __shared__ double sm[];
    const Geo &g = a.g;
    const int TX = zg.TX, TY = zg.TY, KC = zg.KC;
    const int RW = TX + 2, RN = RW * (TY + 2);
    const unsigned blk = cell_block();
    const int ntx = g.nx / TX, nty = g.ny / TY;
    const int i0 = (int)(blk % (unsigned)ntx) * TX, j0 = (int)((blk / (unsigned)ntx) % (unsigned)nty) * TY;
    const int k0 = (int)(blk / (unsigned)(ntx * nty)) * KC;
    const int tid = (int)threadIdx.x, bd = (int)block_dim_x(), tx = tid % TX, ty = tid / TX;
    const int mx = g.nx - 1, my = g.ny - 1, mz = g.nz - 1;
    const int i = i0 + tx, j = j0 + ty;
    const double *f[NF] = {a.nu, a.cc, a.rho, a.Pi};
    auto gidx = [&](int x, int y, int z) { return (long)(z & mz) * g.nxy + (long)(y & my) * g.nx + (x & mx); };
    auto T = [&](int slot, int q, int x, int y) -> double { return sm[(slot * NF + q) * RN + (y + 1) * RW + (x + 1)]; };
    auto load_tile = [&](int kz, int slot) {
      for (int e = tid; e < RN; e += bd) {
        const long o = gidx(i0 + e % RW - 1, j0 + e / RW - 1, kz);
#pragma unroll
        for (int q = 0; q < NF; ++q) sm[(slot * NF + q) * RN + e] = f[q][o];
      }
    };
    double zb[NF];  // the plane below, at the cell
    {
      const long o = gidx(i, j, k0 - 1);
#pragma unroll
      for (int q = 0; q < NF; ++q) zb[q] = f[q][o];
    }
    load_tile(k0, 0);
    load_tile(k0 + 1, 1);
    __syncthreads();
    const double mean = a.rhs_sum[0] / (double)g.n;
    double part = 0.0;
    int cur = 0;
    for (int kk = k0; kk < k0 + KC; ++kk) {
      const int nxt = cur ^ 1;
      const long c = gidx(i, j, kk);
      const double nuc = T(cur, 0, tx, ty);
      const double muc = T(cur, 1, tx, ty) * nuc;
      const double rc = T(cur, 2, tx, ty), pc = T(cur, 3, tx, ty);
      double dmu = 0.0, dpi = 0.0, lrho = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        // field q at the low (m) and high (p) neighbour along d
        auto vm = [&](int q) { return d == 0 ? T(cur, q, tx - 1, ty) : (d == 1 ? T(cur, q, tx, ty - 1) : zb[q]); };
        auto vp = [&](int q) { return d == 0 ? T(cur, q, tx + 1, ty) : (d == 1 ? T(cur, q, tx, ty + 1) : T(nxt, q, tx, ty)); };
        const double id = inv_d(g, d);
        const double c0 = rcp_nr((rc + vm(2)) * 0.5), c1 = rcp_nr((vp(2) + rc) * 0.5);
        const double mum = vm(1) * vm(0), mup = vp(1) * vp(0);
        const double gm0 = (muc - mum) * id, gm1 = (mup - muc) * id;
        const double gp0 = (pc - vm(3)) * id, gp1 = (vp(3) - pc) * id;
        dmu += (c1 * gm1 - c0 * gm0) * id;
        dpi += (c1 * gp1 - c0 * gp0) * id;
        lrho += 0.5 * ((c0 * c0) * (gm0 * gp0) + (c1 * c1) * (gm1 * gp1));
      }
      const double rr = dpi - (a.rhs[c] - mean);
      const double lr = (a.first ? 0.0 : a.lrhs[c]) - muc;
      a.lrhs[c] = lr;
      a.lcc[c] = (a.first ? 0.0 : a.lcc[c]) + nuc * rr;
      a.lp[c] = a.lp[c] + dmu;
      a.lrho[c] = a.lrho[c] + lrho;
      part += lr;
      // this plane becomes the plane below of the next
#pragma unroll
      for (int q = 0; q < NF; ++q) zb[q] = T(cur, q, tx, ty);
      __syncthreads();
      if (kk + 1 < k0 + KC) load_tile(kk + 2, cur);
      __syncthreads();
      cur = nxt;
    }
    if (a.last) block_sum_bd(part, a.partials);
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
    const long c = (long)cell_block() * block_dim_x() + threadIdx.x;
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
    const long c = (long)cell_block() * block_dim_x() + threadIdx.x;
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

static long launch_adj_pois_it(tpx_ctx *c, const PoisArgs &pa);

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
  long nbz = 0;
  for (int i = s.iters - 1; i >= 0; --i) {
    pa.Pi = s.P[i];
    pa.first = i == s.iters - 1;
    pa.last = i == 0;
    launch_spectral_apply(c, ga.lp, ga.nu, ga.scal);
    nbz = launch_adj_pois_it(c, pa);
  }
  launch_finish_sum(c, (int)nbz, 0, SC_ADJ_SUM);  // sum of lam_RHS_Pois from the last reversed iteration
  // G6 and G5, with the Poisson coefficient's share of lam_F' (cc(C), C = 1/F')
  {
    RhsPArgs a{};
    a.g = c->g;
    a.rho = s.rho_new;
    for (int d = 0; d < 3; ++d) { a.ru_new[d] = s.ru_new[d]; a.lnew[d] = ga.lnew[d]; a.lF[d] = ga.lF[d]; }
    a.lrhs = ga.lrhs_p;
    a.cc = ga.cc;
    a.lcc = ga.lcc;
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
      a.lF[d] = ga.lF[d];
    }
    a.lr = ga.lr; a.lrnp1 = ga.lrnp1; a.lrnew = ga.lrnew; a.ld = ga.ld;
    a.c_post = k.dt * k.post;
    a.c_pre = k.dt * k.pre;
    a.last = last;
    tpx_launch<k_adj_rk>(nb, kBlock, 0, c->stream, a);
    TPX_LAUNCH_CHECK();
  }
  return ga.lrhs;
}


// the z-march tile of the adjoint stencils: 32 x 8 cells (all of x and more
// rows when nx < 32), 32 planes a group, halved while the launch has fewer
// than 1024 groups (down to 2); *nblocks the launch's groups
static AdjMomGeom adj_zm_geom(const tpx_ctx *c, long *nblocks, const char *what) {
  const Geo &g = c->g;
  AdjMomGeom zg;
  zg.TX = g.nx < 32 ? g.nx : 32;
  zg.TY = kBlock / zg.TX < g.ny ? kBlock / zg.TX : g.ny;
  zg.KC = g.nz < 32 ? g.nz : 32;
  const long cols = (long)(g.nx / zg.TX) * (g.ny / zg.TY);
  while (zg.KC > 2 && zg.KC % 2 == 0 && cols * (g.nz / zg.KC) < 1024) zg.KC /= 2;
  const int threads = zg.TX * zg.TY;
  *nblocks = cols * (g.nz / zg.KC);
  if ((zg.TX + 2) * (zg.TY + 2) > 2 * threads || (threads & (threads - 1)) != 0 || *nblocks > c->n_partials)
    throw std::logic_error(std::string(what) + ": the grid does not fit the z-march tile");
  return zg;
}

// one reversed Poisson iteration (k_adj_pois_zm); returns its groups (the
// number of lam_RHS block sums it leaves when pa.last is set)
static long launch_adj_pois_it(tpx_ctx *c, const PoisArgs &pa) {
  long nbz = 0;
  const AdjMomGeom zg = adj_zm_geom(c, &nbz, "Poisson adjoint");
  const size_t sh = sizeof(double) * 2 * 4 * (size_t)((zg.TX + 2) * (zg.TY + 2));
  tpx_launch<k_adj_pois_zm>((unsigned)nbz, zg.TX * zg.TY, sh, c->stream, pa, zg);
  TPX_LAUNCH_CHECK();
  return nbz;
}

void adjoint_gas_substep_b(tpx_ctx *c, GasAdjoint &ga, const GasSubstep &s, const Coef &k) {
  const long n = c->g.n;
  const unsigned nb = cells(c);
  MomArgs m{};
  m.g = c->g;
  for (int d = 0; d < 3; ++d) {
    m.lrhs[d] = ga.lrhs[d];
    m.rui[d] = s.ru_int[d];
    m.lup[d] = s.lup ? s.lup + (long)d * n : nullptr;
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
  long nbm = 0;
  const AdjMomGeom zg = adj_zm_geom(c, &nbm, "momentum adjoint");
  const size_t sh = sizeof(double) * 2 * 11 * (size_t)((zg.TX + 2) * (zg.TY + 2));
  tpx_launch<k_adj_mom>((unsigned)nbm, zg.TX * zg.TY, sh, c->stream, m, zg);
  TPX_LAUNCH_CHECK();
  tpx_launch<k_adj_acc>(1, 1024, 0, c->stream, (const double *)c->partials, nbm, c->scal + SC_ADJ_A);
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
  launch_adj_pois_it(c, pa);
  tpx_launch<k_adj_pois_c>(nb, kBlock, 0, c->stream, pa);
  TPX_LAUNCH_CHECK();
  tpx_launch<k_adj_fold_add>(nb, kBlock, 0, c->stream, c->g, (const double *)lF[0], (const double *)lF[1],
                             (const double *)lF[2], atr);
  TPX_LAUNCH_CHECK();
}

}  // namespace tpx
