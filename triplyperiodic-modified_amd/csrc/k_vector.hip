// k_vector.hip -- the passive vector of the reference (grid.cpp:616-715):
// a momentum-like field RV at the faces, transported by the gas velocity
// U = RU_int/Rho_face, diffused with eta, forced by V_Source (or by the gas's
// x velocity on its y component), and projected with its own pressure Q
// (a Poisson solve with the constant density Rho_forV).  One thread per
// cell; every product and sum in the order of the reference's whole-array
// passes (built with -ffp-contract=off), so the stencil parts round as the
// reference's do.
#include <math.h>

#include "tpx_internal.h"

namespace tpx {

namespace {

struct RU3c { const double *__restrict__ a[3]; };
struct RU3w { double *__restrict__ a[3]; };

struct IxV {
  int mx, my, mz, nx;
  long nxy;
  __device__ __forceinline__ long operator()(int i, int j, int k) const {
    return (long)(k & mz) * nxy + (long)(j & my) * nx + (i & mx);
  }
};

__device__ __forceinline__ IxV ixv(const Geo &g) { return IxV{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy}; }

__device__ __forceinline__ bool cell(const Geo &g, int &i, int &j, int &k, long &c) {
  c = (long)xcd_block(blockIdx.x, gridDim.x) * block_dim_x() + threadIdx.x;
  if (c >= g.n) return false;
  i = (int)(c % g.nx);
  const long r = c / g.nx;
  j = (int)(r % g.ny);
  k = (int)(r / g.ny);
  return true;
}

__device__ __forceinline__ double bsum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = block_dim_x() / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

struct RvArgs {
  const double *rv_int[3], *rv_base[3];
  double *rv_np1[3], *rv_new[3];
  const double *ru_int[3], *rho_int;
  const double *filt;    // S2_type != 0: the filtered RU_int.x / Rho_face.x
  const double *scal;    // SC_RV_S2: the type-0 source summed over the cells
  double eta, rho_forv, c_post, c_pre, ncell;
  int s2_type, rk3;
};

// grid::Update_RV_WOQ (grid.cpp:646-692) at the three faces of one cell
struct k_rv_woq {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const Geo &g, const RvArgs &a) {
    int i, j, k;
    long c;
    if (!cell(g, i, j, k, c)) return;
    const IxV ix = ixv(g);
    const double id[3] = {g.idx, g.idy, g.idz}, id2[3] = {g.idx2, g.idy2, g.idz2};
    const int e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    auto at = [&](int di, int dj, int dk) { return ix(i + di, j + dj, k + dk); };
    // V = RV_int / Rho_forV (tensor0.cpp:143-147)
    auto V = [&](int b, long q) { return a.rv_int[b][q] / a.rho_forv; };
    // div_F2C(V) at the cell offset by (di, dj, dk) (tensor0.cpp:249-256)
    auto divV = [&](int di, int dj, int dk) {
      const long q = at(di, dj, dk);
      return (V(0, at(di + 1, dj, dk)) - V(0, q)) * g.idx + (V(1, at(di, dj + 1, dk)) - V(1, q)) * g.idy +
             (V(2, at(di, dj, dk + 1)) - V(2, q)) * g.idz;
    };
    // U.d = RU_int.d / Rho_face.d, Rho_face = I_C2F(Rho_int) (grid.cpp:868)
    auto U = [&](int d, long q, int di, int dj, int dk) {
      const long qm = at(di - e[d][0], dj - e[d][1], dk - e[d][2]);
      return a.ru_int[d][q] / ((a.rho_int[q] + a.rho_int[qm]) * 0.5);
    };
    const double div0 = divV(0, 0, 0);
    for (int ax = 0; ax < 3; ++ax) {
      const int* ea = e[ax];
      const double g1 = (div0 - divV(-ea[0], -ea[1], -ea[2])) * id[ax];  // grad_C2F (tensor1.cpp:259-270)
      double lap = 0.0;                                                  // del2 (tensor0.cpp:258-266)
      for (int d = 0; d < 3; ++d) {
        const double l = (V(ax, at(e[d][0], e[d][1], e[d][2])) - 2 * V(ax, c) + V(ax, at(-e[d][0], -e[d][1], -e[d][2]))) *
                         id2[d];
        lap = d == 0 ? l : lap + l;
      }
      double rhs = a.eta / 3. * g1 + a.eta * lap;
      // convection: div_F2C of I_d(RV_int.ax) * I_ax(U.d) (grid.cpp:655-673)
      double conv = 0.0;
      for (int d = 0; d < 3; ++d) {
        const int *ed = e[d];
        // the two faces of direction d around the cell: q0 = the cell, q1 = the cell + e_d
        double T[2];
        for (int s = 0; s < 2; ++s) {
          const int oi = s * ed[0], oj = s * ed[1], ok = s * ed[2];
          const long q = at(oi, oj, ok);
          const double irv = (a.rv_int[ax][q] + a.rv_int[ax][at(oi - ed[0], oj - ed[1], ok - ed[2])]) * 0.5;
          const double iu = (U(d, q, oi, oj, ok) +
                             U(d, at(oi - ea[0], oj - ea[1], ok - ea[2]), oi - ea[0], oj - ea[1], ok - ea[2])) * 0.5;
          T[s] = irv * iu;
        }
        const double t = (T[1] - T[0]) * id[d];
        conv = d == 0 ? t : conv + t;
      }
      rhs -= conv;
      if (a.s2_type == 0) {
        // V_Source (grid.cpp:616-641, vector_source.h): cos(X) on the y faces,
        // zero on the others, each minus its mean over the cells
        const double X = i * g.dx + g.dx / 2., A = 1., K = 1.;
        const double s2 = ax == 1 ? A * cos(K * X) - a.scal[SC_RV_S2] / a.ncell : 0.0;
        rhs += s2;
      } else if (ax == 1) {
        rhs += -1 * a.filt[c];
      }
      const double np1 = a.rv_np1[ax][c] + a.c_post * rhs;
      a.rv_np1[ax][c] = np1;
      a.rv_new[ax][c] = a.rk3 ? np1 : 1 * a.rv_base[ax][c] + a.c_pre * rhs;
    }
  }
};

// The type-1 source's tensor0::Equal_Iy_C2F(dummy.x) on dummy.x itself
// (grid.cpp:683-685): the reference's loop (tensor0.cpp:308-315) runs j
// upwards and reads row j-1 after overwriting it, so along each (i, k) line
// f~_j = (f_j + f~_{j-1}) / 2, with f~_{-1} the old value of the last row (the
// periodic ghost).  f = RU_int.x / Rho_face.x.  One thread per line.
struct k_rv_filter {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const Geo &g, const double *__restrict__ rux,
                                             const double *__restrict__ rho, double *__restrict__ out) {
    const long line = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (line >= (long)g.nx * g.nz) return;
    const int i = (int)(line % g.nx), k = (int)(line / g.nx);
    const IxV ix = ixv(g);
    auto f = [&](int j) {
      const long q = ix(i, j, k);
      return rux[q] / ((rho[q] + rho[ix(i - 1, j, k)]) * 0.5);
    };
    double prev = f(g.ny - 1);
    for (int j = 0; j < g.ny; ++j) {
      const double v = (f(j) + prev) * 0.5;
      out[ix(i, j, k)] = v;
      prev = v;
    }
  }
};

// grid::Compute_RHS_Pois_Q (grid.cpp:694-701) before its mean is removed:
// div_F2C(RV_new / Rho_forV) / (dt pre), with block sums for the mean (and
// for the statistics the sum of squares of V = RV_new / Rho_forV)
struct k_rhs_pois_q {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const Geo &g, RU3c rv, double rho_forv, double inv_dtpre,
                                             double *__restrict__ rhs, double *__restrict__ partials, long nb,
                                             int want_vv) {
    __shared__ double sh[256];
    int i, j, k;
    long c;
    double r = 0.0, vv = 0.0;
    if (cell(g, i, j, k, c)) {
      const IxV ix = ixv(g);
      const double v0 = rv.a[0][c] / rho_forv, v1 = rv.a[1][c] / rho_forv, v2 = rv.a[2][c] / rho_forv;
      r = (rv.a[0][ix(i + 1, j, k)] / rho_forv - v0) * g.idx + (rv.a[1][ix(i, j + 1, k)] / rho_forv - v1) * g.idy +
          (rv.a[2][ix(i, j, k + 1)] / rho_forv - v2) * g.idz;
      r = r * inv_dtpre;
      rhs[c] = r;
      vv = v0 * v0 + v1 * v1 + v2 * v2;
    }
    const double s = bsum(r, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
    if (want_vv) {
      __syncthreads();
      const double t = bsum(vv, sh);
      if (threadIdx.x == 0) partials[nb + blockIdx.x] = t;
    }
  }
};

// grid::Update_RV_WQ (grid.cpp:709-715)
struct k_rv_wq {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const Geo &g, const double *__restrict__ Q, RU3w np1, RU3w nw,
                                             double c_post, double c_pre, int rk3) {
    int i, j, k;
    long c;
    if (!cell(g, i, j, k, c)) return;
    const IxV ix = ixv(g);
    const double gq[3] = {(Q[c] - Q[ix(i - 1, j, k)]) * g.idx, (Q[c] - Q[ix(i, j - 1, k)]) * g.idy,
                          (Q[c] - Q[ix(i, j, k - 1)]) * g.idz};
    for (int d = 0; d < 3; ++d) {
      const double v = np1.a[d][c] + -(c_post)*gq[d];
      np1.a[d][c] = v;
      nw.a[d][c] = rk3 ? v : nw.a[d][c] + -(c_pre)*gq[d];
    }
  }
};

// block sums of the type-0 source's y component cos(X) over the cells
struct k_rv_source_sum {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const Geo &g, double *partials) {
    __shared__ double sh[256];
    int i, j, k;
    long c;
    double v = 0.0;
    if (cell(g, i, j, k, c)) {
      const double X = i * g.dx + g.dx / 2., A = 1., K = 1.;
      v = A * cos(K * X);
    }
    const double s = bsum(v, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
};

unsigned vgrid(const tpx_ctx *c) { return grid1d(c->g.n, 256); }

}  // namespace

void launch_rv_woq(tpx_ctx *c, const Coef &k, const double *const ru_int[3], const double *rho_int) {
  RvArgs a{};
  for (int d = 0; d < 3; ++d) {
    a.rv_int[d] = c->rv[c->rv_int][d];
    a.rv_base[d] = c->rv[c->rv_base][d];
    a.rv_np1[d] = c->rv[c->rv_np1][d];
    a.rv_new[d] = c->rv[c->rv_new][d];
    a.ru_int[d] = ru_int[d];
  }
  a.rho_int = rho_int;
  a.scal = c->scal;
  a.eta = c->p.eta;
  a.rho_forv = c->p.rho_forv;
  a.c_post = k.dt * k.post;
  a.c_pre = k.dt * k.pre;
  a.ncell = (double)c->g.n;
  a.s2_type = c->p.s2_type;
  a.rk3 = k.rk == 3;
  timing_begin(c, TPX_K_OTHER);
  if (a.s2_type != 0) {
    tpx_launch<k_rv_filter>(grid1d((long)c->g.nx * c->g.nz, 256), 256, 0, c->stream, c->g, ru_int[0], rho_int,
                            c->rv_filt);
    TPX_LAUNCH_CHECK();
    a.filt = c->rv_filt;
  }
  tpx_launch<k_rv_woq>(vgrid(c), 256, 0, c->stream, c->g, a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_OTHER);
}

void launch_rhs_pois_q(tpx_ctx *c, const Coef &k, bool want_vv) {
  RU3c rv;
  for (int d = 0; d < 3; ++d) rv.a[d] = c->rv[c->rv_new][d];
  const unsigned nb = vgrid(c);
  timing_begin(c, TPX_K_OTHER);
  tpx_launch<k_rhs_pois_q>(nb, 256, 0, c->stream, c->g, rv, c->p.rho_forv, (1. / (k.dt * k.pre)), c->rhs_q,
                           c->partials, (long)nb, want_vv ? 1 : 0);
  TPX_LAUNCH_CHECK();
  launch_finish_sum(c, (int)nb, 0, SC_SUM_RHS_Q);
  if (want_vv) launch_finish_sum(c, (int)nb, (int)nb, SC_ST_VV);
  timing_end(c, TPX_K_OTHER);
}

void launch_rv_wq(tpx_ctx *c, const Coef &k) {
  RU3w np1, nw;
  for (int d = 0; d < 3; ++d) {
    np1.a[d] = c->rv[c->rv_np1][d];
    nw.a[d] = c->rv[c->rv_new][d];
  }
  timing_begin(c, TPX_K_OTHER);
  tpx_launch<k_rv_wq>(vgrid(c), 256, 0, c->stream, c->g, (const double *)c->Qv, np1, nw, k.dt * k.post,
                      k.dt * k.pre, k.rk == 3 ? 1 : 0);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_OTHER);
}

void launch_rv_source_sum(tpx_ctx *c) {
  const unsigned nb = vgrid(c);
  tpx_launch<k_rv_source_sum>(nb, 256, 0, c->stream, c->g, c->partials);
  TPX_LAUNCH_CHECK();
  launch_finish_sum(c, (int)nb, 0, SC_RV_S2);
}

}  // namespace tpx
