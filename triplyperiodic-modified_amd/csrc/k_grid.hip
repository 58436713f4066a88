// k_grid.hip -- staggered-grid kernels of the RK4 substep (gfx950).
//
// One thread per cell (i,j,k); it owns the cell centre and the three faces
// i, j, k of that cell (x face at x = i dx, etc.).  Periodic neighbours are
// formed by masking indices (grid sizes are powers of two).  Every
// expression keeps the operand order of the reference operator it fuses, and
// the library is built without FP contraction, so a fused kernel rounds
// exactly like the reference's sequence of whole-array passes.
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

// cell (i,j,k) handled by this thread, with the XCD-contiguous block order
__device__ __forceinline__ bool cell_of(const Geo &g, int &i, int &j, int &k, long &c) {
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  c = (long)b * block_dim_x() + threadIdx.x;
  if (c >= g.n) return false;
  i = (int)(c % g.nx);
  long r = c / g.nx;
  j = (int)(r % g.ny);
  k = (int)(r / g.ny);
  return true;
}

// deterministic block sum (fixed tree); result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = block_dim_x() / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

__device__ __forceinline__ double block_max(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = block_dim_x() / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  return sh[0];
}

// U = RU / Rho_face on face a of cell (i,j,k): Equal_Divide(RU, Rho_face)
// with Rho_face = I_C2F(Rho) (tensor1.cpp:278-283, tensor0.cpp:288-296)
template <int A>
__device__ __forceinline__ double Uface(const Ix &ix, const double *__restrict__ ru,
                                        const double *__restrict__ rho, int i, int j, int k) {
  long p = ix(i, j, k);
  long q = ix(i - (A == 0), j - (A == 1), k - (A == 2));
  return ru[p] / ((rho[p] + rho[q]) * 0.5);
}

struct RU3 { const double *__restrict__ a[3]; };

// div_F2C(U) at cell centre (i,j,k) (tensor0.cpp:249-256)
__device__ __forceinline__ double divU(const Geo &g, const Ix &ix, const RU3 &ru,
                                       const double *__restrict__ rho, int i, int j, int k) {
  return (Uface<0>(ix, ru.a[0], rho, i + 1, j, k) - Uface<0>(ix, ru.a[0], rho, i, j, k)) * g.idx +
         (Uface<1>(ix, ru.a[1], rho, i, j + 1, k) - Uface<1>(ix, ru.a[1], rho, i, j, k)) * g.idy +
         (Uface<2>(ix, ru.a[2], rho, i, j, k + 1) - Uface<2>(ix, ru.a[2], rho, i, j, k)) * g.idz;
}

struct RhsArgs {
  RU3 ru_int;
  const double *rho_int;
  const double *ru_base[3];
  const double *ru_np1_in[3];
  const double *S[3];
  double *ru_np1[3];
  double *ru_new[3];
  const double *rho_base;
  const double *rho_np1_in;
  double *rho_np1;
  double *rho_new;
  double c_np1, c_new, c_rho_np1, c_rho_new;
  __device__ void to_global() {
    for (int c = 0; c < 3; ++c) {
      ru_int.a[c] = gptr(ru_int.a[c]);
      ru_base[c] = gptr(ru_base[c]);
      ru_np1_in[c] = gptr(ru_np1_in[c]);
      S[c] = gptr(S[c]);
      ru_np1[c] = gptr(ru_np1[c]);
      ru_new[c] = gptr(ru_new[c]);
    }
    rho_int = gptr(rho_int);
    rho_base = gptr(rho_base);
    rho_np1_in = gptr(rho_np1_in);
    rho_np1 = gptr(rho_np1);
    rho_new = gptr(rho_new);
  }
};

// ---------------------------------------------------------------------------
// K1: Update_RU_WOP (grid.cpp:776-814) fused with Update_Rho
// (grid.cpp:608-614), which reads the same RU_int, as a z-marching tile.
// The right-hand side of momentum component A on face A of cell (i,j,k) is
// (grid.cpp:780-809)
//   -(mp/Vcell) S_A + (mu/3) d_A(div U) + mu lap U_A
//   - sum_d d_d^{F2C}[ I_d^{C2F}(U_A) I_A^{C2F}(RU_d) ] + A RU_A + (Rho_face - Rho0) g_A
// with U = RU_int / Rho_face.  A block owns a TX x TY column of cells and
// walks KC planes along z.  Per plane it stages RU_int and Rho_int (planes
// k and k+1, with a halo of 2 below and 1 above in x and y) in LDS, forms
// every face velocity of the plane once (U_z also for plane k+1) and the
// cell divergences of U once, and evaluates the stencil from those values
// in the reference's operand order: each U costs one division instead of
// the ~19 per face component of a per-cell evaluation, and every value read
// is the same double, so the results are bit-identical.  Values of plane
// k-1 that the stencil still needs are carried in registers; the next plane
// is prefetched into registers while the current one is computed.
// ---------------------------------------------------------------------------
struct ZmGeom { int TX, TY, KC; };

template <bool WITH_RHO, bool HAS_S, bool RK3>
struct k_rhs_ru {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const Geo &g, const Coef &kc, const RhsArgs &a, const ZmGeom &zg) {
#include "k_rhs_ru_body.inc"
  }
};

// The single-box launch keeps the body in the kernel itself: inlined from a
// device function it allocates scalar registers worse (measured 10 % slower;
// for a batch, whose kernels read their arguments from a table, the same
// kernel-resident body measured the same as the struct form).
template <bool WITH_RHO, bool HAS_S, bool RK3>
__global__ __launch_bounds__(256) void k_rhs_ru_kernel(Geo g, Coef kc, RhsArgs a, ZmGeom zg) {
#include "k_rhs_ru_body.inc"
}

// G1 alone: Update_Rho (grid.cpp:608-614)
template <bool RK3>
struct k_update_rho {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const RhsArgs &a) {
  int i, j, k;
  long c;
  if (!cell_of(g, i, j, k, c)) return;
  Ix ix = make_ix(g);
  const double *__restrict__ u0 = a.ru_int.a[0];
  const double *__restrict__ u1 = a.ru_int.a[1];
  const double *__restrict__ u2 = a.ru_int.a[2];
  double R = (u0[ix(i + 1, j, k)] - u0[c]) * g.idx + (u1[ix(i, j + 1, k)] - u1[c]) * g.idy +
             (u2[ix(i, j, k + 1)] - u2[c]) * g.idz;
  double np1 = a.rho_np1_in[c] + a.c_rho_np1 * R;
  a.rho_np1[c] = np1;
  if (!RK3) a.rho_new[c] = 1 * a.rho_base[c] + a.c_rho_new * R;
}
};

struct DivArgs {
  RU3 ru_new;
  const double *rho_new;
  const double *Q;
  double *divergence;
  double *rhs;
  double *partials;
  const double *scal;
  double inv_dtpre;
  double ncell;
  __device__ void to_global() {
    for (int q = 0; q < 3; ++q) ru_new.a[q] = gptr(ru_new.a[q]);
    rho_new = gptr(rho_new);
    Q = gptr(Q);
    divergence = gptr(divergence);
    rhs = gptr(rhs);
    partials = gptr(partials);
    scal = gptr(scal);
  }
};

// (k/Cp) del2(1/Rho_new) - R/(Cp P0_new Vcell) Q' (grid.cpp:832-839)
__device__ __forceinline__ double low_mach_div(const Geo &g, const Ix &ix, const Coef &kc,
                                               const DivArgs &a, int i, int j, int k, long c) {
  const double *__restrict__ rho = a.rho_new;
  double fc = 1. / rho[c];
  double lap = ((1. / rho[ix(i + 1, j, k)]) - 2 * fc + (1. / rho[ix(i - 1, j, k)])) * g.idx2 +
               ((1. / rho[ix(i, j + 1, k)]) - 2 * fc + (1. / rho[ix(i, j - 1, k)])) * g.idy2 +
               ((1. / rho[ix(i, j, k + 1)]) - 2 * fc + (1. / rho[ix(i, j, k - 1)])) * g.idz2;
  double q = a.Q ? a.Q[c] : 0.0;
  if (kc.cooling && a.Q) q -= a.scal[SC_SUM_Q] / a.ncell;  // make_mean_zero
  double P0n = a.scal[SC_P0_NEW];
  return kc.k_over_cp * lap + -kc.R / (kc.cp * P0n * kc.Vcell) * q;
}

// G5+G6: Compute_Div_U_new + Compute_RHS_Pois (grid.cpp:832-849).  MODE 0
// fuses both, MODE 1 stores the divergence field, MODE 2 reads it.  The raw
// right-hand side is stored; its mean is removed where it is consumed.
template <int MODE>
struct k_div_rhs {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const Coef &kc, const DivArgs &a) {
  __shared__ double sh[kBlock];
  int i, j, k;
  long c;
  bool ok = cell_of(g, i, j, k, c);
  Ix ix = make_ix(g);
  double v = 0.0;
  if (ok) {
    double dv;
    if (MODE == 2) dv = a.divergence[c];
    else dv = low_mach_div(g, ix, kc, a, i, j, k, c);
    if (MODE == 1) {
      a.divergence[c] = dv;
    } else {
      const double *__restrict__ rho = a.rho_new;
      double du = (Uface<0>(ix, a.ru_new.a[0], rho, i + 1, j, k) - Uface<0>(ix, a.ru_new.a[0], rho, i, j, k)) * g.idx +
                  (Uface<1>(ix, a.ru_new.a[1], rho, i, j + 1, k) - Uface<1>(ix, a.ru_new.a[1], rho, i, j, k)) * g.idy +
                  (Uface<2>(ix, a.ru_new.a[2], rho, i, j, k + 1) - Uface<2>(ix, a.ru_new.a[2], rho, i, j, k)) * g.idz;
      v = (du - dv) * a.inv_dtpre;
      a.rhs[c] = v;
    }
  }
  if (MODE != 1) {
    double s = block_sum(v, sh);
    if (threadIdx.x == 0) a.partials[xcd_block(blockIdx.x, gridDim.x)] = s;
  }
}
};

struct ProjArgs {
  const double *P;
  double *np1[3];
  double *nw[3];
  double *partials;
  long npart;
  double c_np1, c_new;
  __device__ void to_global() {
    P = gptr(P);
    for (int q = 0; q < 3; ++q) np1[q] = gptr(np1[q]);
    for (int q = 0; q < 3; ++q) nw[q] = gptr(nw[q]);
    partials = gptr(partials);
  }
};

// G10: Update_RU_WP (grid.cpp:857-863); at RK4_count==3 RU_new aliases
// RU_np1 and the block sums of RU_np1 feed make_mean_U0.
template <bool RK3>
struct k_project {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const ProjArgs &a) {
  __shared__ double sh[kBlock];
  int i, j, k;
  long c;
  bool ok = cell_of(g, i, j, k, c);
  Ix ix = make_ix(g);
  double s[3] = {0, 0, 0};
  if (ok) {
    const double *__restrict__ P = a.P;
    double pc = P[c];
    double gr[3] = {(pc - P[ix(i - 1, j, k)]) * g.idx, (pc - P[ix(i, j - 1, k)]) * g.idy,
                    (pc - P[ix(i, j, k - 1)]) * g.idz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double v = a.np1[d][c] + a.c_np1 * gr[d];
      a.np1[d][c] = v;
      s[d] = v;
      if (!RK3) a.nw[d][c] = a.nw[d][c] + a.c_new * gr[d];
    }
  }
  if (RK3) {
    unsigned b = xcd_block(blockIdx.x, gridDim.x);
    for (int d = 0; d < 3; ++d) {
      double t = block_sum(s[d], sh);
      if (threadIdx.x == 0) a.partials[d * a.npart + b] = t;
      __syncthreads();
    }
  }
}
};

// tensor0::make_mean_U0 (tensor0.cpp:386-396): out = (in - mean) + a
struct k_shift_mean {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(long n, const double *__restrict__ in,
                                                       double *__restrict__ out, const double *scal,
                                                       int slot, double ncell, double a) {
  long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
  if (c >= n) return;
  double mean = scal[slot] / ncell;  // mean() = sum()/size_tot()
  out[c] = (in[c] - mean) + a;
}
};

// generic reduction of one field into per-block partials
template <int OP>
struct k_reduce {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(long n, const double *__restrict__ f,
                                                   double *partials) {
  __shared__ double sh[kBlock];
  long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
  double v = 0.0;
  if (c < n) v = OP == 0 ? f[c] : (OP == 2 ? f[c] * f[c] : fabs(f[c]));
  double s = OP == 1 ? block_max(v, sh) : block_sum(v, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
};

// sum (OP 0) or max (OP 1) of nparts partials into scal[slot], fixed order:
// thread t combines partials t, t + 1024, t + 2048, ... (coalesced reads,
// four independent accumulators), then the block tree
template <int OP>
struct k_finish {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(const double *__restrict__ partials, long nparts,
                                                 double *scal, int slot) {
  __shared__ double sh[1024];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  long q = threadIdx.x;
  for (; q + 3 * 1024 < nparts; q += 4 * 1024) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double x = partials[q + u * 1024];
      acc[u] = OP == 0 ? acc[u] + x : fmax(acc[u], x);
    }
  }
  for (int u = 0; q < nparts; q += 1024, ++u) acc[u] = OP == 0 ? acc[u] + partials[q] : fmax(acc[u], partials[q]);
  double v = OP == 0 ? (acc[0] + acc[1]) + (acc[2] + acc[3]) : fmax(fmax(acc[0], acc[1]), fmax(acc[2], acc[3]));
  double s = OP == 0 ? block_sum(v, sh) : block_max(v, sh);
  if (threadIdx.x == 0) scal[slot] = s;
}
};

// grid::Update_P0 (grid.cpp:816-830) on the device scalars
struct k_update_p0 {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(const Coef &kc, double *scal, double ncell, double c_np1, double c_new) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double met = scal[SC_SUM_Q] / ncell;
  scal[SC_MET] = met;
  double dP0 = kc.cooling ? 0.0 : -kc.R / kc.cv * met / kc.Vcell;
  scal[SC_DP0_DT] = dP0;
  scal[SC_P0_NP1] += c_np1 * dP0;
  if (kc.rk != 3) scal[SC_P0_NEW] = scal[SC_P0] + c_new * dP0;
  else scal[SC_P0_NEW] = scal[SC_P0_NP1];
}
};

struct k_scalar_copy {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(double *scal, int dst, int src) {
  if (threadIdx.x == 0 && blockIdx.x == 0) scal[dst] = scal[src];
}
};

struct k_fill {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(double *f, double v, long n) {
  long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
  if (c < n) f[c] = v;
}
};

struct k_copy {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(double *__restrict__ d, const double *__restrict__ s, long n) {
  long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
  if (c < n) d[c] = s[c];
}
};

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned cells_grid(const tpx_ctx *c) { return grid1d(c->g.n, kBlock); }

void launch_update_rho(tpx_ctx *c, const Coef &k, const double *const ru_int[3],
                       const double *rho_base, const double *rho_np1_in, double *rho_np1,
                       double *rho_new) {
  RhsArgs a{};
  for (int d = 0; d < 3; ++d) a.ru_int.a[d] = ru_int[d];
  a.rho_base = rho_base; a.rho_np1_in = rho_np1_in; a.rho_np1 = rho_np1; a.rho_new = rho_new;
  a.c_rho_np1 = -(k.dt * k.post);
  a.c_rho_new = -k.dt * k.pre;
  timing_begin(c, TPX_K_GRID_RHS);
  if (k.rk == 3) tpx_launch<k_update_rho<true>>(cells_grid(c), kBlock, 0, c->stream, c->g, a);
  else tpx_launch<k_update_rho<false>>(cells_grid(c), kBlock, 0, c->stream, c->g, a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_GRID_RHS);
}

void launch_rhs_ru(tpx_ctx *c, const Coef &k, bool with_rho, const double *const ru_int[3],
                   const double *rho_int, const double *const ru_base[3],
                   const double *const ru_np1_in[3], const double *const S[3],
                   double *const ru_np1[3], double *const ru_new[3], const double *rho_base,
                   const double *rho_np1_in, double *rho_np1, double *rho_new) {
  RhsArgs a{};
  for (int d = 0; d < 3; ++d) {
    a.ru_int.a[d] = ru_int[d];
    a.ru_base[d] = ru_base[d];
    a.ru_np1_in[d] = ru_np1_in[d];
    a.S[d] = S ? S[d] : nullptr;
    a.ru_np1[d] = ru_np1[d];
    a.ru_new[d] = ru_new[d];
  }
  a.rho_int = rho_int;
  a.rho_base = rho_base; a.rho_np1_in = rho_np1_in; a.rho_np1 = rho_np1; a.rho_new = rho_new;
  a.c_np1 = k.dt * k.post;
  a.c_new = k.dt * k.pre;
  a.c_rho_np1 = -(k.dt * k.post);
  a.c_rho_new = -k.dt * k.pre;
  bool hs = S != nullptr, r3 = k.rk == 3;
  const Geo &g = c->g;
  ZmGeom zg;
  zg.TX = g.nx < 32 ? g.nx : 32;
  zg.TY = 256 / zg.TX < g.ny ? 256 / zg.TX : g.ny;
  // planes per block: 32, halved while the launch (over every box of a
  // batch) has fewer than 1024 blocks, down to 2.  Each block re-reads two
  // halo planes of the four raw fields, so a short march costs bandwidth, but
  // a small launch needs the blocks: one 64^3 box runs 128 -> 20 us at 2
  // planes, one 128^3 box 170 -> 84 us at 8; 64 boxes of 64^3 keep 32.
  zg.KC = g.nz < 32 ? g.nz : 32;
  {
    const long cols = (long)(g.nx / zg.TX) * (g.ny / zg.TY) * (g_rec ? g_rec->nbox : 1u);
    while (zg.KC > 2 && zg.KC % 2 == 0 && cols * (g.nz / zg.KC) < 1024) zg.KC /= 2;
  }
  unsigned nb = (unsigned)((g.nx / zg.TX) * (g.ny / zg.TY) * (g.nz / zg.KC));
  int threads = zg.TX * zg.TY;
  if ((zg.TX + 3) * (zg.TY + 3) > 2 * threads) throw std::logic_error("z-march tile exceeds two points per thread");
  size_t sh = sizeof(double) * (size_t)(8 * (zg.TX + 3) * (zg.TY + 3) + 4 * (zg.TX + 2) * (zg.TY + 2) +
                                        (zg.TX + 1) * (zg.TY + 1));
  timing_begin(c, TPX_K_GRID_RHS);
#define TPX_K1(WR, HS, R3)                                                            \
  do {                                                                                \
    if (g_rec) tpx_launch<k_rhs_ru<WR, HS, R3>>(nb, threads, sh, c->stream, g, k, a, zg); \
    else k_rhs_ru_kernel<WR, HS, R3><<<nb, threads, sh, c->stream>>>(g, k, a, zg);         \
  } while (0)
  if (with_rho) {
    if (hs) { if (r3) TPX_K1(true, true, true); else TPX_K1(true, true, false); }
    else { if (r3) TPX_K1(true, false, true); else TPX_K1(true, false, false); }
  } else {
    if (hs) { if (r3) TPX_K1(false, true, true); else TPX_K1(false, true, false); }
    else { if (r3) TPX_K1(false, false, true); else TPX_K1(false, false, false); }
  }
#undef TPX_K1
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_GRID_RHS);
}

// grid::Update_Passive_Scalar (grid.cpp:711-745) in one pass: the scalar on
// the faces is I_C2F of its int value (the reference keeps it from the
// previous substep's new value, grid.cpp:744, 476), U = RU_int / I_C2F(Rho_int),
// flux-form transport, molecular diffusion D_M del2, the source (form 0:
// C_Source (grid.cpp:686-708, scalar_source.h:4-8: cos(X) at the cell
// centre) minus its mean, held in SC_PS_S1 as a sum; form 1: -Ix_F2C(U.x)),
// then the RK4 updates of np1 (in place) and new, operand for operand.
struct PsArgs {
  RU3 ru_int;
  const double *rho_int, *cint, *cbase;
  double *cnp1, *cnew;
  const double *scal;
  double d_m, c_post, c_pre, ncell;
  int s1_type, rk3;
  __device__ void to_global() {
    for (int q = 0; q < 3; ++q) ru_int.a[q] = gptr(ru_int.a[q]);
    rho_int = gptr(rho_int);
    cint = gptr(cint);
    cbase = gptr(cbase);
    cnp1 = gptr(cnp1);
    cnew = gptr(cnew);
    scal = gptr(scal);
  }
};

struct k_passive_scalar {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const PsArgs &a) {
  int i, j, k;
  long c;
  if (!cell_of(g, i, j, k, c)) return;
  Ix ix = make_ix(g);
  const double *__restrict__ rho = a.rho_int;
  const double *__restrict__ C = a.cint;
  const long m[3] = {ix(i - 1, j, k), ix(i, j - 1, k), ix(i, j, k - 1)};
  const long p[3] = {ix(i + 1, j, k), ix(i, j + 1, k), ix(i, j, k + 1)};
  const double id[3] = {g.idx, g.idy, g.idz}, id2[3] = {g.idx2, g.idy2, g.idz2};
  const double rc = rho[c], cc = C[c];
  double div = 0.0, lap = 0.0, ux0 = 0.0, ux1 = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double u0 = a.ru_int.a[d][c] / ((rc + rho[m[d]]) * 0.5);
    const double u1 = a.ru_int.a[d][p[d]] / ((rho[p[d]] + rc) * 0.5);
    const double f0 = u0 * ((cc + C[m[d]]) * 0.5), f1 = u1 * ((C[p[d]] + cc) * 0.5);
    div = d == 0 ? (f1 - f0) * id[d] : div + (f1 - f0) * id[d];
    const double l = (C[p[d]] - 2 * cc + C[m[d]]) * id2[d];
    lap = d == 0 ? l : lap + l;
    if (d == 0) { ux0 = u0; ux1 = u1; }
  }
  double rhs = a.d_m * lap + -1 * div;
  if (a.s1_type == 0) {
    const double X = g.dx / 2. + i * g.dx, A = 1., kk = 1.;
    rhs += A * cos(kk * X) - a.scal[SC_PS_S1] / a.ncell;
  } else {
    rhs += -1 * ((ux0 + ux1) * 0.5);
  }
  const double np1 = a.cnp1[c] + a.c_post * rhs;
  a.cnp1[c] = np1;
  a.cnew[c] = a.rk3 ? np1 : 1 * a.cbase[c] + a.c_pre * rhs;
}
};

// block sums of C_Source over the cells (its mean is removed in form 0)
struct k_ps_source_sum {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, double *partials) {
  __shared__ double sh[kBlock];
  int i, j, k;
  long c;
  double v = 0.0;
  if (cell_of(g, i, j, k, c)) {
    const double X = g.dx / 2. + i * g.dx, A = 1., kk = 1.;
    v = A * cos(kk * X);
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}
};

void launch_div_rhs(tpx_ctx *c, const Coef &k, int mode, const double *const ru_new[3],
                    const double *rho_new, const double *Q, double *divergence, double *rhs) {
  DivArgs a{};
  for (int d = 0; d < 3; ++d) a.ru_new.a[d] = ru_new[d];
  a.rho_new = rho_new; a.Q = Q; a.divergence = divergence; a.rhs = rhs;
  a.partials = c->partials; a.scal = c->scal;
  a.inv_dtpre = (1. / (k.dt * k.pre));
  a.ncell = (double)c->g.n;
  unsigned nb = cells_grid(c);
  timing_begin(c, TPX_K_DIV_RHS);
  if (mode == 0) tpx_launch<k_div_rhs<0>>(nb, kBlock, 0, c->stream, c->g, k, a);
  else if (mode == 1) tpx_launch<k_div_rhs<1>>(nb, kBlock, 0, c->stream, c->g, k, a);
  else tpx_launch<k_div_rhs<2>>(nb, kBlock, 0, c->stream, c->g, k, a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_DIV_RHS);
  if (mode != 1) {
    timing_begin(c, TPX_K_OTHER);
    tpx_launch<k_finish<0>>(1, 1024, 0, c->stream, c->partials, nb, c->scal, SC_SUM_RHS);
    TPX_LAUNCH_CHECK();
    timing_end(c, TPX_K_OTHER);
  }
}

void launch_project(tpx_ctx *c, const Coef &k, const double *P, double *const ru_np1[3],
                    double *const ru_new[3], bool copy_np1_to_new, bool want_sums) {
  ProjArgs a{};
  a.P = P;
  for (int d = 0; d < 3; ++d) { a.np1[d] = ru_np1[d]; a.nw[d] = ru_new[d]; }
  a.partials = c->partials;
  unsigned nb = cells_grid(c);
  a.npart = nb;
  a.c_np1 = -(k.dt * k.post);
  a.c_new = -(k.dt * k.pre);
  timing_begin(c, TPX_K_PROJECT);
  if (copy_np1_to_new) tpx_launch<k_project<true>>(nb, kBlock, 0, c->stream, c->g, a);
  else tpx_launch<k_project<false>>(nb, kBlock, 0, c->stream, c->g, a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_PROJECT);
  if (copy_np1_to_new && want_sums) {
    timing_begin(c, TPX_K_OTHER);
    for (int d = 0; d < 3; ++d)
      tpx_launch<k_finish<0>>(1, 1024, 0, c->stream, c->partials + (long)d * nb, nb, c->scal, SC_MEAN_RU0 + d);
    TPX_LAUNCH_CHECK();
    timing_end(c, TPX_K_OTHER);
  }
}

void launch_shift_mean(tpx_ctx *c, const double *const in[3], double *const out[3],
                       const double a[3]) {
  timing_begin(c, TPX_K_OTHER);
  for (int d = 0; d < 3; ++d)
    tpx_launch<k_shift_mean>(cells_grid(c), kBlock, 0, c->stream, c->g.n, in[d], out[d], c->scal,
                                                         SC_MEAN_RU0 + d, (double)c->g.n, a[d]);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_OTHER);
}

void launch_reduce(tpx_ctx *c, const double *f, int op, int slot) {
  unsigned nb = cells_grid(c);
  timing_begin(c, TPX_K_OTHER);
  if (op == 0) {
    tpx_launch<k_reduce<0>>(nb, kBlock, 0, c->stream, c->g.n, f, c->partials);
    tpx_launch<k_finish<0>>(1, 1024, 0, c->stream, c->partials, nb, c->scal, slot);
  } else if (op == 2) {  // sum of squares
    tpx_launch<k_reduce<2>>(nb, kBlock, 0, c->stream, c->g.n, f, c->partials);
    tpx_launch<k_finish<0>>(1, 1024, 0, c->stream, c->partials, nb, c->scal, slot);
  } else {
    tpx_launch<k_reduce<1>>(nb, kBlock, 0, c->stream, c->g.n, f, c->partials);
    tpx_launch<k_finish<1>>(1, 1024, 0, c->stream, c->partials, nb, c->scal, slot);
  }
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_OTHER);
}

void launch_finish_sum(tpx_ctx *c, int nparts, int part_offset, int slot) {
  tpx_launch<k_finish<0>>(1, 1024, 0, c->stream, c->partials + part_offset, nparts, c->scal, slot);
  TPX_LAUNCH_CHECK();
}

void launch_passive_scalar(tpx_ctx *c, const Coef &k, const double *const ru_int[3], const double *rho_int) {
  PsArgs a{};
  for (int d = 0; d < 3; ++d) a.ru_int.a[d] = ru_int[d];
  a.rho_int = rho_int;
  a.cint = c->ps[c->ps_int];
  a.cbase = c->ps[c->ps_base];
  a.cnp1 = c->ps[c->ps_np1];
  a.cnew = c->ps[c->ps_new];
  a.scal = c->scal;
  a.d_m = c->p.d_m;
  a.c_post = k.dt * k.post;
  a.c_pre = k.dt * k.pre;
  a.ncell = (double)c->g.n;
  a.s1_type = c->p.s1_type;
  a.rk3 = k.rk == 3;
  timing_begin(c, TPX_K_OTHER);
  tpx_launch<k_passive_scalar>(cells_grid(c), kBlock, 0, c->stream, c->g, a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_OTHER);
}

void launch_passive_scalar_source_sum(tpx_ctx *c) {
  const unsigned nb = cells_grid(c);
  tpx_launch<k_ps_source_sum>(nb, kBlock, 0, c->stream, c->g, c->partials);
  TPX_LAUNCH_CHECK();
  launch_finish_sum(c, (int)nb, 0, SC_PS_S1);
}

void launch_update_p0(tpx_ctx *c, const Coef &k) {
  timing_begin(c, TPX_K_OTHER);
  tpx_launch<k_update_p0>(1, 64, 0, c->stream, k, c->scal, (double)c->g.n, k.dt * k.post, k.dt * k.pre);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_OTHER);
}

void launch_scalar_copy(tpx_ctx *c, int dst, int src) {
  tpx_launch<k_scalar_copy>(1, 64, 0, c->stream, c->scal, dst, src);
  TPX_LAUNCH_CHECK();
}

void launch_fill(tpx_ctx *c, double *f, double v, long n) {
  tpx_launch<k_fill>(grid1d(n, kBlock), kBlock, 0, c->stream, f, v, n);
  TPX_LAUNCH_CHECK();
}

void launch_copy(tpx_ctx *c, double *dst, const double *src, long n) {
  tpx_launch<k_copy>(grid1d(n, kBlock), kBlock, 0, c->stream, dst, src, n);
  TPX_LAUNCH_CHECK();
}

}  // namespace tpx
