// k_stats.hip -- grid::Statistics (grid.cpp:902-982) as device reductions.
//
// At a step boundary the statistics read Rho, RU (= RU_np1) with
// U = RU / I_C2F(Rho), the particles' u and T, and three things the last
// substep left behind in the reference: the concentration C deposited at that
// substep's positions (Update_Particle), its pre-projection velocity
// U = RU_new / I_C2F(Rho_new) (Compute_RHS_Pois writes U before the solve; the
// statistics take its mean square before recomputing U) and the gathered gas
// x velocity part.ug.  The last three are captured during the fourth substep
// when the statistics are on (stats_capture_*); stats_reduce computes the
// rest.  Reductions run per block into partial arrays, then one block per
// quantity finishes them; only scalars go to the host.
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

// op: 0 sum, 1 max, 2 min
__device__ __forceinline__ double combine(int op, double a, double b) {
  return op == 0 ? a + b : (op == 1 ? fmax(a, b) : fmin(a, b));
}
__device__ __forceinline__ double identity(int op) { return op == 0 ? 0.0 : (op == 1 ? -INFINITY : INFINITY); }

__device__ __forceinline__ double block_reduce(int op, double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = combine(op, sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// cell quantities: 0 sum Rho, 1 sum P0/(R Rho), 2-4 sum U_c RU_c, 5 sum C,
// 6 max Rho, 7 min Rho, 8 max C, 9 min C, 10-12 max U_c
constexpr int kCellStats = 13;
constexpr int kCellOps[kCellStats] = {0, 0, 0, 0, 0, 0, 1, 2, 1, 2, 1, 1, 1};

struct CellStatArgs {
  Geo g;
  const double *rho, *ru[3], *C, *scal;
  double R;
  double *part;   // kCellStats x stride
  long stride;
};

struct k_stat_cells {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const CellStatArgs &a) {
    __shared__ double sh[kBlock];
    const Geo &g = a.g;
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    double v[kCellStats];
#pragma unroll
    for (int q = 0; q < kCellStats; ++q) v[q] = identity(kCellOps[q]);
    if (c < g.n) {
      Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
      const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / g.nxy);
      const double r = a.rho[c];
      const double tg = (a.scal[SC_P0] / a.R) / r;
      const long m[3] = {ix(i - 1, j, k), ix(i, j - 1, k), ix(i, j, k - 1)};
      v[0] = r;
      v[1] = tg;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double ru = a.ru[d][c];
        const double u = ru / ((r + a.rho[m[d]]) * 0.5);
        v[2 + d] = u * ru;
        v[10 + d] = u;
      }
      const double cv = a.C[c];
      v[5] = cv;
      v[6] = r;
      v[7] = r;
      v[8] = cv;
      v[9] = cv;
    }
#pragma unroll
    for (int q = 0; q < kCellStats; ++q) {
      const double t = block_reduce(kCellOps[q], v[q], sh);
      if (threadIdx.x == 0) a.part[q * a.stride + blockIdx.x] = t;
    }
  }
};

// particle quantities: 0 max u, 1 max T, 2 sum T
constexpr int kPartStats = 3;
constexpr int kPartOps[kPartStats] = {1, 1, 0};

struct k_stat_particles {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const double *u, const double *T, long np, double *part, long stride) {
    __shared__ double sh[kBlock];
    const long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    double v[kPartStats] = {-INFINITY, -INFINITY, 0.0};
    if (p < np) {
      v[0] = u[p];
      v[1] = T[p];
      v[2] = T[p];
    }
#pragma unroll
    for (int q = 0; q < kPartStats; ++q) {
      const double t = block_reduce(kPartOps[q], v[q], sh);
      if (threadIdx.x == 0) part[q * stride + blockIdx.x] = t;
    }
  }
};

// sum over the faces of (RU_new_c / I_C2F(Rho_new)_c)^2
struct k_stat_tke2 {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const double *ru0, const double *ru1, const double *ru2,
                                             const double *rho, double *part) {
    __shared__ double sh[kBlock];
    const long c = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    double v = 0.0;
    if (c < g.n) {
      Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
      const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / g.nxy);
      const double r = rho[c];
      const double u0 = ru0[c] / ((r + rho[ix(i - 1, j, k)]) * 0.5);
      const double u1 = ru1[c] / ((r + rho[ix(i, j - 1, k)]) * 0.5);
      const double u2 = ru2[c] / ((r + rho[ix(i, j, k - 1)]) * 0.5);
      v = u0 * u0 + u1 * u1 + u2 * u2;
    }
    const double t = block_reduce(0, v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
};

// the gathered gas x velocity at the particles (particle.cpp:340-360), max
struct k_stat_ug {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const double *x, const double *y, const double *z,
                                             const double *ru0, const double *rho, long np, double *part) {
    __shared__ double sh[kBlock];
    const long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    double v = -INFINITY;
    if (p < np) {
      Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
      const int ii = (int)(x[p] * g.idx), jj = (int)(y[p] * g.idy), kk = (int)(z[p] * g.idz);
      const double xl = (x[p] * g.idx - ((int)(x[p] * g.idx)));
      const long c0 = ix(ii, jj, kk), c1 = ix(ii + 1, jj, kk);
      const double u1 = ru0[c1] / ((rho[c1] + rho[c0]) * 0.5);
      const double u0 = ru0[c0] / ((rho[c0] + rho[ix(ii - 1, jj, kk)]) * 0.5);
      v = xl * u1 + (1 - xl) * u0;
    }
    const double t = block_reduce(1, v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
};

// one block per quantity: out[q] = reduction (op q) of part[q][0..nparts)
struct k_stat_finish {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const double *part, long stride, long nparts, unsigned ops, double *out) {
    __shared__ double sh[kBlock];
    const int q = blockIdx.x;
    const int op = (int)((ops >> (2 * q)) & 3u);
    double v = identity(op);
    for (long b = threadIdx.x; b < nparts; b += kBlock) v = combine(op, v, part[q * stride + b]);
    const double t = block_reduce(op, v, sh);
    if (threadIdx.x == 0) out[q] = t;
  }
};

void check_launch(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

double *stats_buffer(tpx_ctx *c) {
  if (!c->st_part) {
    // partial sums: 16 arrays of one value per block of cells or particles,
    // then the results
    const long nb = std::max((long)grid1d(c->g.n, kBlock), (long)grid1d(std::max<int64_t>(c->pt.cap, 1), kBlock));
    c->st_stride = nb;
    if (hipMalloc((void **)&c->st_part, sizeof(double) * (16 * (size_t)nb + 32)) != hipSuccess) throw std::bad_alloc();
  }
  return c->st_part;
}

}  // namespace

void stats_capture_particles(tpx_ctx *c, const double *const ru_int[3], const double *rho_int) {
  Particles &pt = c->pt;
  if (pt.n == 0) {
    launch_fill(c, c->C, 0.0, c->g.n);
    launch_fill(c, c->scal + SC_ST_UG, -1e9, 1);  // particle::max with no particles
    return;
  }
  // the concentration at this substep's positions, with their bins
  if (!(c->p.twc || c->p.nu * TPX_PI * c->p.dp * c->p.k != 0.0)) launch_sort_stage(c, TPX_INT, true);
  launch_concentration(c);
  double *part = stats_buffer(c);
  const int s = pt.int_buf();
  const unsigned nb = grid1d(pt.n, kBlock);
  tpx_launch<k_stat_ug>(nb, kBlock, 0, c->stream, c->g, (const double *)pt.q[0][s], (const double *)pt.q[1][s],
                        (const double *)pt.q[2][s], ru_int[0], rho_int, (long)pt.n, part);
  check_launch("k_stat_ug");
  tpx_launch<k_stat_finish>(1, kBlock, 0, c->stream, (const double *)part, c->st_stride, (long)nb, 1u,
                            c->scal + SC_ST_UG);
  check_launch("k_stat_finish");
}

void stats_capture_velocity(tpx_ctx *c, const double *const ru_new[3], const double *rho_new) {
  double *part = stats_buffer(c);
  const unsigned nb = grid1d(c->g.n, kBlock);
  tpx_launch<k_stat_tke2>(nb, kBlock, 0, c->stream, c->g, ru_new[0], ru_new[1], ru_new[2], rho_new, part);
  check_launch("k_stat_tke2");
  tpx_launch<k_stat_finish>(1, kBlock, 0, c->stream, (const double *)part, c->st_stride, (long)nb, 0u,
                            c->scal + SC_ST_TKE2);
  check_launch("k_stat_finish");
}

// raw[0..12]: the cell reductions, raw[13..15]: the particle ones
// particle::trajectory (particle.cpp:717-745): the position of the particle
// whose flag is m*stride, for m < count, into out[3m..3m+2] (left as filled
// when no particle has that flag)
struct k_track {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const double *__restrict__ x, const double *__restrict__ y,
                                             const double *__restrict__ z, const int32_t *__restrict__ flag, long n,
                                             int stride, int count, double *__restrict__ out) {
    const long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (p >= n) return;
    const int f = flag[p];
    if (f < 0 || f % stride != 0 || f / stride >= count) return;
    const int m = f / stride;
    out[3 * m] = x[p];
    out[3 * m + 1] = y[p];
    out[3 * m + 2] = z[p];
  }
};

void launch_track(tpx_ctx *c, int stride, int count, double *out) {
  launch_fill(c, out, NAN, 3L * count);
  Particles &pt = c->pt;
  if (pt.n > 0) {
    tpx_launch<k_track>(grid1d(pt.n, kBlock), kBlock, 0, c->stream, (const double *)pt.q[0][pt.st_base],
                        (const double *)pt.q[1][pt.st_base], (const double *)pt.q[2][pt.st_base],
                        (const int32_t *)pt.flag, (long)pt.n, stride, count, out);
    check_launch("k_track");
  }
}

void stats_reduce(tpx_ctx *c, const double *rho, const double *const ru[3], double raw[16]) {
  double *part = stats_buffer(c);
  double *out = part + 16 * c->st_stride;
  const unsigned nbc = grid1d(c->g.n, kBlock);
  CellStatArgs a{};
  a.g = c->g;
  a.rho = rho;
  for (int d = 0; d < 3; ++d) a.ru[d] = ru[d];
  a.C = c->C;
  a.scal = c->scal;
  a.R = c->R;
  a.part = part;
  a.stride = c->st_stride;
  tpx_launch<k_stat_cells>(nbc, kBlock, 0, c->stream, a);
  check_launch("k_stat_cells");
  unsigned ops = 0;
  for (int q = 0; q < kCellStats; ++q) ops |= (unsigned)kCellOps[q] << (2 * q);
  tpx_launch<k_stat_finish>(kCellStats, kBlock, 0, c->stream, (const double *)part, c->st_stride, (long)nbc, ops, out);
  check_launch("k_stat_finish");
  Particles &pt = c->pt;
  if (pt.n > 0) {
    const unsigned nbp = grid1d(pt.n, kBlock);
    tpx_launch<k_stat_particles>(nbp, kBlock, 0, c->stream, (const double *)pt.q[3][pt.st_base],
                                 (const double *)pt.q[6][pt.st_base], (long)pt.n, part, c->st_stride);
    check_launch("k_stat_particles");
    unsigned pops = 0;
    for (int q = 0; q < kPartStats; ++q) pops |= (unsigned)kPartOps[q] << (2 * q);
    tpx_launch<k_stat_finish>(kPartStats, kBlock, 0, c->stream, (const double *)part, c->st_stride, (long)nbp, pops,
                              out + kCellStats);
    check_launch("k_stat_finish");
  } else {
    launch_fill(c, out + kCellStats, -1e9, 2);  // particle::max with no particles
    launch_fill(c, out + kCellStats + 2, 0.0, 1);
  }
  if (hipMemcpyAsync(raw, out, sizeof(double) * 16, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    throw std::runtime_error("statistics download");
}

}  // namespace tpx
