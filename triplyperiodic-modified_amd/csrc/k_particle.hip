// k_particle.hip -- Lagrangian particle path of the RK4 substep (gfx950).
//
// Particles are SoA arrays kept physically ordered by the cell that holds
// them at the start of each step.  Every substep they are re-binned by the
// floor cell of x_new with a stable, deterministic counting sort:
//   count (integer atomics) -> exclusive scan -> placement -> per-cell
//   insertion sort of the placed indices,
// which yields a CSR (cell -> [start, end) of a permutation).  Deposits are
// then pull-form segment reductions: each grid point sums the particles of
// the few cells that can reach it, in a fixed order, so no floating-point
// atomics are used and every run rounds identically.
//   velocity deposit (particle.cpp:494-520): face i of cell c sums the
//     (1 - x_l) shares of cell c and the x_l shares of cell c - e_x;
//   heat / concentration deposit (particle.cpp:522-555,620-654): centre c
//     sums the 27 cells around it, keeping the particles whose rounded
//     index puts c among their 8 cloud-in-cell targets.
#include <math.h>

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

struct AdvArgs {
  Geo g;
  const double *ru[3];      // RU_int
  const double *rho;        // Rho_int
  const double *base[7];
  const double *in_int[7];
  const double *np1_in[7];
  double *np1[7];
  double *nw[7];
  double *base_w[3];        // base x,y,z (shifted by the wrap)
  const double *rhs_t;
  const double *tau;
  const double *eps;
  double *contrib;
  int32_t *key;
  long n;
  double c_post, c_pre;     // dt*post, dt*pre
  double iTp, g3[3];        // 1/Tp, ParticleGravity*g
  double iS, Dp, I0;        // 1/(mp Cvp), particle diameter, lamp intensity
  double epsp;              // scalar epsilonp
  double L[3];
  int rk3;
};

// P1-P4, P6 shares and P7 wrap for one particle (particle.cpp:314-399,
// 477-492, 494-520, 58-104)
__global__ __launch_bounds__(kBlock) void k_advance(AdvArgs a) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const Geo &g = a.g;
  Ix ix = make_ix(g);
  double xi = a.in_int[0][p], yi = a.in_int[1][p], zi = a.in_int[2][p];
  // gas2part_velocity (particle.cpp:340-360): x2i and the fractional part
  int ii = (int)(xi * g.idx), jj = (int)(yi * g.idy), kk = (int)(zi * g.idz);
  double xl = (xi * g.idx - ((int)(xi * g.idx)));
  double yl = (yi * g.idy - ((int)(yi * g.idy)));
  double zl = (zi * g.idz - ((int)(zi * g.idz)));
  const double *__restrict__ rho = a.rho;
  const double *__restrict__ r0 = a.ru[0];
  const double *__restrict__ r1 = a.ru[1];
  const double *__restrict__ r2 = a.ru[2];
  long c000 = ix(ii, jj, kk);
  // U = RU_int / Rho_face on the faces the particle reads
  double ux1 = r0[ix(ii + 1, jj, kk)] / ((rho[ix(ii + 1, jj, kk)] + rho[c000]) * 0.5);
  double ux0 = r0[c000] / ((rho[c000] + rho[ix(ii - 1, jj, kk)]) * 0.5);
  double uy1 = r1[ix(ii, jj + 1, kk)] / ((rho[ix(ii, jj + 1, kk)] + rho[c000]) * 0.5);
  double uy0 = r1[c000] / ((rho[c000] + rho[ix(ii, jj - 1, kk)]) * 0.5);
  double uz1 = r2[ix(ii, jj, kk + 1)] / ((rho[ix(ii, jj, kk + 1)] + rho[c000]) * 0.5);
  double uz0 = r2[c000] / ((rho[c000] + rho[ix(ii, jj, kk - 1)]) * 0.5);
  double gas[3] = {xl * ux1 + (1 - xl) * ux0, yl * uy1 + (1 - yl) * uy0, zl * uz1 + (1 - zl) * uz0};
  double frac[3] = {xl, yl, zl};
  double itp = a.tau ? 1. / a.tau[p] : a.iTp;
  double nv[7], n1[7];
  // update_position (particle.cpp:314-338)
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double ui = a.in_int[3 + d][p];
    n1[d] = a.np1_in[d][p] + a.c_post * ui;
    nv[d] = a.rk3 ? n1[d] : a.base[d][p] + a.c_pre * ui;
  }
  // update_velocity (particle.cpp:362-399) and the deposit shares
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double ui = a.in_int[3 + d][p];
    double rhs = (gas[d] - ui) * itp;
    n1[3 + d] = a.np1_in[3 + d][p] + a.c_post * (rhs + a.g3[d]);
    nv[3 + d] = a.rk3 ? n1[3 + d] : a.base[3 + d][p] + a.c_pre * (rhs + a.g3[d]);
    a.contrib[6 * p + 2 * d] = frac[d] * rhs;            // to face ii+1
    a.contrib[6 * p + 2 * d + 1] = (1 - frac[d]) * rhs;  // to face ii
  }
  // update_Temp (particle.cpp:477-492)
  {
    double epsp = a.eps ? a.eps[p] : a.epsp;
    double alpha = 0.25 * epsp * TPX_PI * a.Dp * a.Dp * a.I0;
    double rt = a.rhs_t[p];
    n1[6] = a.np1_in[6][p] + a.c_post * a.iS * (rt + alpha);
    nv[6] = a.rk3 ? n1[6] : a.base[6][p] + a.c_pre * a.iS * (rt + alpha);
  }
  // periodic wrap of x_new, x and x_np1 (particle.cpp:58-104); x_int keeps its value
  double bx[3] = {a.base[0][p], a.base[1][p], a.base[2][p]};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (nv[d] > a.L[d]) {
      nv[d] -= a.L[d]; bx[d] -= a.L[d]; n1[d] -= a.L[d];
      a.base_w[d][p] = bx[d];
    } else if (nv[d] < 0) {
      nv[d] += a.L[d]; bx[d] += a.L[d]; n1[d] += a.L[d];
      a.base_w[d][p] = bx[d];
    }
  }
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    a.np1[q][p] = n1[q];
    a.nw[q][p] = nv[q];
  }
  int ki = (int)(nv[0] * g.idx), kj = (int)(nv[1] * g.idy), kz = (int)(nv[2] * g.idz);
  a.key[p] = (int32_t)ix(ki, kj, kz);
}

// floor-cell key of one stage of positions
__global__ void k_keys(Geo g, const double *x, const double *y, const double *z, int32_t *key, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  Ix ix = make_ix(g);
  key[p] = (int32_t)ix((int)(x[p] * g.idx), (int)(y[p] * g.idy), (int)(z[p] * g.idz));
}

__global__ void k_zero_i(int32_t *a, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) a[p] = 0;
}

__global__ void k_count(const int32_t *__restrict__ key, int32_t *cnt, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) atomicAdd(&cnt[key[p]], 1);
}

__global__ void k_copy_i(int32_t *__restrict__ d, const int32_t *__restrict__ s, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) d[p] = s[p];
}

__global__ void k_place(const int32_t *__restrict__ key, int32_t *cursor, int32_t *perm, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) {
    int32_t pos = atomicAdd(&cursor[key[p]], 1);
    perm[pos] = (int32_t)p;
  }
}

// per-cell insertion sort of the placed indices: the result is the stable
// order (by physical index) whatever order the placement atomics took
__global__ void k_fixup(const int32_t *__restrict__ start, int32_t *perm, long ncell) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int32_t b = start[c], e = start[c + 1];
  for (int32_t q = b + 1; q < e; ++q) {
    int32_t v = perm[q];
    int32_t r = q - 1;
    while (r >= b && perm[r] > v) { perm[r + 1] = perm[r]; --r; }
    perm[r + 1] = v;
  }
}

__global__ void k_iota(int32_t *p, long n) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = (int32_t)q;
}

// ---- exclusive scan of int32 (three phases, block size 256 x 8) ----
constexpr int kScanPer = 8;
constexpr int kScanChunk = kBlock * kScanPer;

__global__ __launch_bounds__(kBlock) void k_scan_blocks(int32_t *data, long n, int32_t *block_sums) {
  __shared__ int32_t sh[kBlock];
  long base = (long)blockIdx.x * kScanChunk + (long)threadIdx.x * kScanPer;
  int32_t v[kScanPer];
  int32_t tot = 0;
#pragma unroll
  for (int e = 0; e < kScanPer; ++e) {
    long q = base + e;
    v[e] = q < n ? data[q] : 0;
    tot += v[e];
  }
  sh[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int32_t t = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int32_t run = sh[threadIdx.x] - tot;  // exclusive prefix of this thread
#pragma unroll
  for (int e = 0; e < kScanPer; ++e) {
    long q = base + e;
    if (q < n) data[q] = run;
    run += v[e];
  }
  if (threadIdx.x == kBlock - 1) block_sums[blockIdx.x] = sh[kBlock - 1];
}

__global__ __launch_bounds__(kBlock) void k_scan_add(int32_t *data, long n, const int32_t *offs) {
  long q = (long)blockIdx.x * kScanChunk + threadIdx.x;
  int32_t o = offs[blockIdx.x];
  for (int e = 0; e < kScanPer; ++e, q += kBlock)
    if (q < n) data[q] += o;
}

// ---- deposits ----
struct VelDepArgs {
  Geo g;
  const double *contrib;
  const int32_t *perm;
  const int32_t *start;
  double *S[3];
};

// P6 pull: S_a(c) = sum_{p in c} (1-frac_a) RHS_a + sum_{p in c - e_a} frac_a RHS_a
__global__ __launch_bounds__(kBlock) void k_vel_deposit(VelDepArgs a) {
  const Geo &g = a.g;
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  long c = (long)b * blockDim.x + threadIdx.x;
  if (c >= g.n) return;
  Ix ix = make_ix(g);
  int i = (int)(c % g.nx);
  long r = c / g.nx;
  int j = (int)(r % g.ny), k = (int)(r / g.ny);
  const double *__restrict__ cb = a.contrib;
  double s[3] = {0, 0, 0};
  for (int32_t q = a.start[c]; q < a.start[c + 1]; ++q) {
    long p = a.perm[q];
    s[0] += cb[6 * p + 1];
    s[1] += cb[6 * p + 3];
    s[2] += cb[6 * p + 5];
  }
  long nb[3] = {ix(i - 1, j, k), ix(i, j - 1, k), ix(i, j, k - 1)};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    for (int32_t q = a.start[nb[d]]; q < a.start[nb[d] + 1]; ++q) {
      long p = a.perm[q];
      s[d] += cb[6 * p + 2 * d];
    }
    a.S[d][c] = s[d];
  }
}

struct TempArgs {
  Geo g;
  const double *x, *y, *z;
  const double *Tp;          // T_new or T_int
  const double *rho;
  const double *scal;
  int p0_slot;
  double R;
  double iTth;
  double *rhs_t;
  long n;
};

// P8/P10 gather: T = (P0/R)/Rho (grid.cpp:882), cloud-in-cell at the particle
// (particle.cpp:402-456), RHS_T = (Tg - T_p) Nu pi Dp k (particle.cpp:460-474)
__global__ __launch_bounds__(kBlock) void k_temp_gather(TempArgs a) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const Geo &g = a.g;
  Ix ix = make_ix(g);
  double x = a.x[p], y = a.y[p], z = a.z[p];
  int ii = (int)(round(x * g.idx)), jj = (int)(round(y * g.idy)), kk = (int)(round(z * g.idz));
  double xl = x * g.idx - ii + 0.5, yl = y * g.idy - jj + 0.5, zl = z * g.idz - kk + 0.5;
  double at = a.scal[a.p0_slot] / a.R;
  const double *__restrict__ rho = a.rho;
  auto T = [&](int i, int j, int k) { return at / rho[ix(i, j, k)]; };
  double Tg = xl * yl * zl * T(ii, jj, kk) +
              (1 - xl) * yl * zl * T(ii - 1, jj, kk) + xl * (1 - yl) * zl * T(ii, jj - 1, kk) +
              xl * yl * (1 - zl) * T(ii, jj, kk - 1) +
              (1 - xl) * (1 - yl) * zl * T(ii - 1, jj - 1, kk) +
              (1 - xl) * yl * (1 - zl) * T(ii - 1, jj, kk - 1) +
              xl * (yl - 1) * (zl - 1) * T(ii, jj - 1, kk - 1) +
              (1 - xl) * (1 - yl) * (1 - zl) * T(ii - 1, jj - 1, kk - 1);
  a.rhs_t[p] = (Tg - a.Tp[p]) * a.iTth;
}

struct CicDepArgs {
  Geo g;
  const double *x, *y, *z;
  const double *val;         // RHS_T, or nullptr for the concentration (weight 1)
  const int32_t *perm;
  const int32_t *start;
  double *out;
  double *partials;
};

// P5/P9 pull: centre c sums the cloud-in-cell shares of the particles of the
// 27 cells around it (particle.cpp:522-555, 582-654)
__global__ __launch_bounds__(kBlock) void k_cic_deposit(CicDepArgs a) {
  __shared__ double sh[kBlock];
  const Geo &g = a.g;
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  long c = (long)b * blockDim.x + threadIdx.x;
  Ix ix = make_ix(g);
  double acc = 0.0;
  if (c < g.n) {
    int i = (int)(c % g.nx);
    long r = c / g.nx;
    int j = (int)(r % g.ny), k = (int)(r / g.ny);
    for (int dk = -1; dk <= 1; ++dk)
      for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) {
          long f = ix(i + di, j + dj, k + dk);
          for (int32_t q = a.start[f]; q < a.start[f + 1]; ++q) {
            long p = a.perm[q];
            double x = a.x[p], y = a.y[p], z = a.z[p];
            int ii = (int)(round(x * g.idx)), jj = (int)(round(y * g.idy)), kk = (int)(round(z * g.idz));
            // targets are {ii-1, ii} along each axis (periodic)
            bool hx = ((ii & ix.mx) == i), lx = (((ii - 1) & ix.mx) == i);
            bool hy = ((jj & ix.my) == j), ly = (((jj - 1) & ix.my) == j);
            bool hz = ((kk & ix.mz) == k), lz = (((kk - 1) & ix.mz) == k);
            if (!(hx || lx) || !(hy || ly) || !(hz || lz)) continue;
            double xl = x * g.idx - ii + 0.5, yl = y * g.idy - jj + 0.5, zl = z * g.idz - kk + 0.5;
            double wx = hx ? xl : (1 - xl), wy = hy ? yl : (1 - yl), wz = hz ? zl : (1 - zl);
            double w = wx * wy * wz;
            acc += a.val ? w * a.val[p] : w;
          }
        }
    a.out[c] = acc;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && a.partials) a.partials[b] = sh[0];
}

struct PermArgs {
  const int32_t *perm;
  const double *src[7];
  double *dst_a[7];
  double *dst_b[7];
  const double *s1[3];
  double *d1[3];
  const int32_t *flag_s;
  int32_t *flag_d;
  long n;
};

// physical re-sort: gather the base state into cell order; the int stage
// (equal to the base at a step boundary) gets the same values
__global__ __launch_bounds__(kBlock) void k_permute(PermArgs a) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  long s = a.perm[q];
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    double v = a.src[d][s];
    a.dst_a[d][q] = v;
    a.dst_b[d][q] = v;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (a.s1[d]) a.d1[d][q] = a.s1[d][s];
  a.flag_d[q] = a.flag_s[s];
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static void check(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

void launch_iota(tpx_ctx *c, int32_t *p, int64_t n) {
  if (n <= 0) return;
  k_iota<<<grid1d(n, kBlock), kBlock, 0, c->stream>>>(p, n);
  TPX_LAUNCH_CHECK();
}

void launch_exclusive_scan(tpx_ctx *c, int32_t *data, int64_t n) {
  // recursive three-phase scan; block sums of level l live in scan_tmp
  std::vector<std::pair<int32_t *, long>> levels;
  int32_t *cur = data;
  long len = n;
  int32_t *tmp = c->scan_tmp;
  while (true) {
    long nb = (len + kScanChunk - 1) / kScanChunk;
    k_scan_blocks<<<(unsigned)nb, kBlock, 0, c->stream>>>(cur, len, tmp);
    TPX_LAUNCH_CHECK();
    levels.push_back({cur, len});
    if (nb == 1) break;
    cur = tmp;
    len = nb;
    tmp = tmp + nb;
  }
  // propagate the block offsets downwards
  for (int l = (int)levels.size() - 2; l >= 0; --l) {
    long len_l = levels[l].second;
    long nb = (len_l + kScanChunk - 1) / kScanChunk;
    k_scan_add<<<(unsigned)nb, kBlock, 0, c->stream>>>(levels[l].first, len_l, levels[l + 1].first);
    TPX_LAUNCH_CHECK();
  }
}

// stable counting sort of the particles by key; writes perm / start
static void counting_sort(tpx_ctx *c, const int32_t *key, int32_t *perm, int32_t *start) {
  Particles &pt = c->pt;
  long n = pt.n, nc = c->g.n;
  k_zero_i<<<grid1d(nc + 1, kBlock), kBlock, 0, c->stream>>>(pt.cursor, nc + 1);
  if (n > 0) k_count<<<grid1d(n, kBlock), kBlock, 0, c->stream>>>(key, pt.cursor, n);
  TPX_LAUNCH_CHECK();
  launch_exclusive_scan(c, pt.cursor, nc + 1);
  k_copy_i<<<grid1d(nc + 1, kBlock), kBlock, 0, c->stream>>>(start, pt.cursor, nc + 1);
  if (n > 0) k_place<<<grid1d(n, kBlock), kBlock, 0, c->stream>>>(key, pt.cursor, perm, n);
  k_fixup<<<grid1d(nc, kBlock), kBlock, 0, c->stream>>>(start, perm, nc);
  TPX_LAUNCH_CHECK();
}

void launch_particle_advance(tpx_ctx *c, const Coef &k, const double *const ru_int[3],
                             const double *rho_int) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  AdvArgs a{};
  a.g = c->g;
  for (int d = 0; d < 3; ++d) a.ru[d] = ru_int[d];
  a.rho = rho_int;
  int sb = pt.st_base, si = pt.st_int;
  int sn = pt.np1_is_base ? pt.st_base : pt.st_np1;
  for (int q = 0; q < 7; ++q) {
    a.base[q] = pt.q[q][sb];
    a.in_int[q] = pt.q[q][si];
    a.np1_in[q] = pt.q[q][sn];
    a.np1[q] = pt.q[q][pt.st_np1];
    a.nw[q] = pt.q[q][pt.st_new];
  }
  for (int d = 0; d < 3; ++d) a.base_w[d] = pt.q[d][sb];
  a.rhs_t = pt.rhs_t;
  a.tau = pt.has_tau ? pt.tau : nullptr;
  a.eps = pt.has_eps ? pt.eps : nullptr;
  a.contrib = pt.contrib;
  a.key = pt.key;
  a.n = pt.n;
  a.c_post = k.dt * k.post;
  a.c_pre = k.dt * k.pre;
  a.iTp = c->Tp != 0 ? 1. / c->Tp : 0.;
  for (int d = 0; d < 3; ++d) a.g3[d] = (double)c->p.particle_gravity * c->gv[d];
  a.iS = 1. / (c->mp * c->p.cvp);
  a.Dp = c->p.dp;
  a.I0 = c->I0;
  a.epsp = c->p.epsilonp;
  a.L[0] = c->p.lx; a.L[1] = c->p.ly; a.L[2] = c->p.lz;
  a.rk3 = k.rk == 3;
  timing_begin(c, TPX_K_PARTICLE);
  k_advance<<<grid1d(pt.n, kBlock), kBlock, 0, c->stream>>>(a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_PARTICLE);
  pt.np1_is_base = false;  // the np1 arrays now hold x_np1 = x + dt*post*u_int
}

void launch_velocity_deposit(tpx_ctx *c) {
  Particles &pt = c->pt;
  VelDepArgs a{};
  a.g = c->g;
  a.contrib = pt.contrib;
  a.perm = pt.perm_int;
  a.start = pt.start_int;
  for (int d = 0; d < 3; ++d) a.S[d] = c->S[d];
  timing_begin(c, TPX_K_DEPOSIT);
  k_vel_deposit<<<grid1d(c->g.n, kBlock), kBlock, 0, c->stream>>>(a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_DEPOSIT);
}

void launch_sort_new(tpx_ctx *c) {
  Particles &pt = c->pt;
  timing_begin(c, TPX_K_SORT);
  counting_sort(c, pt.key, pt.perm_new, pt.start_new);
  timing_end(c, TPX_K_SORT);
}

// sort the particles by the floor cell of one stage's positions
void launch_sort_stage(tpx_ctx *c, int stage, bool into_int) {
  Particles &pt = c->pt;
  int s = stage == TPX_BASE ? pt.st_base : stage == TPX_INT ? pt.st_int : stage == TPX_NEW ? pt.st_new : pt.st_np1;
  timing_begin(c, TPX_K_SORT);
  if (pt.n > 0) {
    k_keys<<<grid1d(pt.n, kBlock), kBlock, 0, c->stream>>>(c->g, pt.q[0][s], pt.q[1][s], pt.q[2][s], pt.key, pt.n);
    TPX_LAUNCH_CHECK();
  }
  counting_sort(c, pt.key, into_int ? pt.perm_int : pt.perm_new, into_int ? pt.start_int : pt.start_new);
  timing_end(c, TPX_K_SORT);
}

void launch_temp_gather(tpx_ctx *c, int stage, const double *rho, int p0_slot) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  int s = stage == TPX_INT ? pt.st_int : pt.st_new;
  TempArgs a{};
  a.g = c->g;
  a.x = pt.q[0][s]; a.y = pt.q[1][s]; a.z = pt.q[2][s];
  a.Tp = pt.q[6][s];
  a.rho = rho;
  a.scal = c->scal;
  a.p0_slot = p0_slot;
  a.R = c->R;
  a.iTth = c->p.nu * TPX_PI * c->p.dp * c->p.k;
  a.rhs_t = pt.rhs_t;
  a.n = pt.n;
  timing_begin(c, TPX_K_PARTICLE);
  k_temp_gather<<<grid1d(pt.n, kBlock), kBlock, 0, c->stream>>>(a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_PARTICLE);
}

void launch_temp_deposit(tpx_ctx *c, int stage, bool use_sorted_new) {
  Particles &pt = c->pt;
  int s = stage == TPX_INT ? pt.st_int : pt.st_new;
  CicDepArgs a{};
  a.g = c->g;
  a.x = pt.q[0][s]; a.y = pt.q[1][s]; a.z = pt.q[2][s];
  a.val = pt.rhs_t;
  a.perm = use_sorted_new ? pt.perm_new : pt.perm_int;
  a.start = use_sorted_new ? pt.start_new : pt.start_int;
  a.out = c->Q;
  a.partials = c->partials;
  unsigned nb = grid1d(c->g.n, kBlock);
  timing_begin(c, TPX_K_DEPOSIT);
  k_cic_deposit<<<nb, kBlock, 0, c->stream>>>(a);
  TPX_LAUNCH_CHECK();
  launch_finish_sum(c, (int)nb, 0, SC_SUM_Q);
  timing_end(c, TPX_K_DEPOSIT);
}

void launch_concentration(tpx_ctx *c) {
  Particles &pt = c->pt;
  int s = pt.st_int;
  CicDepArgs a{};
  a.g = c->g;
  a.x = pt.q[0][s]; a.y = pt.q[1][s]; a.z = pt.q[2][s];
  a.val = nullptr;
  a.perm = pt.perm_int;
  a.start = pt.start_int;
  a.out = c->C;
  a.partials = c->partials;
  unsigned nb = grid1d(c->g.n, kBlock);
  timing_begin(c, TPX_K_DEPOSIT);
  k_cic_deposit<<<nb, kBlock, 0, c->stream>>>(a);
  TPX_LAUNCH_CHECK();
  launch_finish_sum(c, (int)nb, 0, SC_SUM_C);
  timing_end(c, TPX_K_DEPOSIT);
}

// Physically reorder the base state into the cell order described by
// perm_int (the sort of the current base positions), duplicate it into the
// int stage (equal to the base at a step boundary), and reset perm_int to the
// identity; start_int stays valid.
void launch_particle_permute(tpx_ctx *c) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  int sb = pt.st_base;
  int free_a = pt.st_np1, free_b = pt.st_int;  // np1 is stale (np1_is_base), int == base
  PermArgs a{};
  a.perm = pt.perm_int;
  for (int q = 0; q < 7; ++q) {
    a.src[q] = pt.q[q][sb];
    a.dst_a[q] = pt.q[q][free_a];
    a.dst_b[q] = pt.q[q][free_b];
  }
  double **own[3] = {&pt.rhs_t, &pt.tau, &pt.eps};
  bool use[3] = {true, pt.has_tau, pt.has_eps};
  for (int d = 0; d < 3; ++d) {
    a.s1[d] = use[d] ? *own[d] : nullptr;
    a.d1[d] = pt.spare[d];
  }
  a.flag_s = pt.flag;
  a.flag_d = pt.itmp;
  a.n = pt.n;
  timing_begin(c, TPX_K_SORT);
  k_permute<<<grid1d(pt.n, kBlock), kBlock, 0, c->stream>>>(a);
  TPX_LAUNCH_CHECK();
  for (int d = 0; d < 3; ++d)
    if (use[d]) std::swap(*own[d], pt.spare[d]);
  std::swap(pt.flag, pt.itmp);
  pt.st_base = free_a;
  pt.st_np1 = sb;
  pt.st_int = free_b;
  pt.np1_is_base = true;
  launch_iota(c, pt.perm_int, pt.n);
  timing_end(c, TPX_K_SORT);
  check(hipGetLastError());
}

}  // namespace tpx
