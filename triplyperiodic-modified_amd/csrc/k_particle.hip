// k_particle.hip -- Lagrangian particle path of the RK4 substep (gfx950).
//
// Particles are SoA arrays kept physically ordered by the cell of their base
// position (re-sorted once per step).  Every substep they are re-binned by
// the floor cell of x_new.  Because a particle moves less than one cell per
// substep, the new bins are built without atomics from the 27 base cells
// around each cell: a count pass, an exclusive scan, and a placement pass
// that visits the neighbours in increasing physical index, so the order is
// stable and identical on every run.  A device flag set by the advance
// kernel switches to a general counting sort (integer atomics, then a
// per-cell insertion sort) if any particle moved further.
//
// Deposits are pull-form segment reductions, never floating-point atomics:
//   velocity (particle.cpp:494-520): face i of cell c sums the (1 - x_l)
//     shares of the particles in cell c and the x_l shares of cell c - e_x,
//     read contiguously because the advance kernel writes the shares at the
//     particles' sorted positions;
//   heat (particle.cpp:620-654): the gather kernel runs in sorted order and
//     writes one record per particle (cloud-in-cell fractions, RHS_T and the
//     rounded cell); centre c sums the records of the 27 cells around it
//     whose rounded cell puts c among their 8 targets.
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
  double *contrib;           // shares at sorted positions, 6 arrays (null: no two-way coupling)
  long cstride;
  const int32_t *inv_int;    // particle -> sorted position of x_int
  const int32_t *base_key;   // base cell (null: no move check)
  int32_t *moved;
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
    if (a.contrib) {  // component 2d: share of face ii+1, 2d+1: of face ii
      long q = a.inv_int[p];
      a.contrib[(2 * d) * a.cstride + q] = frac[d] * rhs;
      a.contrib[(2 * d + 1) * a.cstride + q] = (1 - frac[d]) * rhs;
    }
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
  long nk = ix(ki, kj, kz);
  a.key[p] = (int32_t)nk;
  if (a.base_key) {  // has the particle left the 27 cells around its base cell?
    long bk = a.base_key[p];
    int bi = (int)(bk % g.nx), bj = (int)((bk / g.nx) % g.ny), bz = (int)(bk / g.nxy);
    int di = (ki - bi) & ix.mx, dj = (kj - bj) & ix.my, dz = (kz - bz) & ix.mz;
    bool far = (di > 1 && di < ix.mx) || (dj > 1 && dj < ix.my) || (dz > 1 && dz < ix.mz);
    if (far) a.moved[0] = 1;
  }
}

// floor-cell key of one stage of positions
__global__ void k_keys(Geo g, const double *x, const double *y, const double *z, int32_t *key, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  Ix ix = make_ix(g);
  key[p] = (int32_t)ix((int)(x[p] * g.idx), (int)(y[p] * g.idy), (int)(z[p] * g.idz));
}

__global__ void k_set_i(int32_t *a, int32_t v, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) a[p] = v;
}

// ---- general path (gated on *moved != 0) ----
__global__ void k_zero_gated(int32_t *a, long n, const int32_t *moved) {
  if (!moved[0]) return;
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) a[p] = 0;
}

__global__ void k_count(const int32_t *__restrict__ key, int32_t *cnt, long n, const int32_t *moved) {
  if (!moved[0]) return;
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) atomicAdd(&cnt[key[p]], 1);
}

__global__ void k_copy_i(int32_t *__restrict__ d, const int32_t *__restrict__ s, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) d[p] = s[p];
}

__global__ void k_place(const int32_t *__restrict__ key, int32_t *cursor, int32_t *perm, long n,
                        const int32_t *moved) {
  if (!moved[0]) return;
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) {
    int32_t pos = atomicAdd(&cursor[key[p]], 1);
    perm[pos] = (int32_t)p;
  }
}

// per-cell insertion sort of the placed indices: the result is the stable
// order (by physical index) whatever order the placement atomics took
__global__ void k_fixup(const int32_t *__restrict__ start, int32_t *perm, int32_t *inv, long ncell,
                        const int32_t *moved) {
  if (!moved[0]) return;
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int32_t b = start[c], e = start[c + 1];
  for (int32_t q = b + 1; q < e; ++q) {
    int32_t v = perm[q];
    int32_t r = q - 1;
    while (r >= b && perm[r] > v) { perm[r + 1] = perm[r]; --r; }
    perm[r + 1] = v;
  }
  for (int32_t q = b; q < e; ++q) inv[perm[q]] = q;
}

// ---- incremental path (gated on *moved == 0) ----
// order of {i-1, i, i+1} (mod n) by increasing index, as offsets
__device__ __forceinline__ void sorted_offsets(int i, int n, int o[3]) {
  if (i == 0) { o[0] = 0; o[1] = 1; o[2] = -1; }
  else if (i == n - 1) { o[0] = 1; o[1] = -1; o[2] = 0; }
  else { o[0] = -1; o[1] = 0; o[2] = 1; }
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
struct TempArgs {
  Geo g;
  const double *x, *y, *z;
  const double *Tp;          // T_new or T_int
  const double *rho;
  const double *scal;
  int p0_slot;
  double R;
  double iTth;
  const int32_t *perm;       // sorted position -> particle
  double *rhs_t;             // by particle
  double4 *rec;              // by sorted position
  int32_t *rkey;
  long n;
};

// P8/P10 gather in sorted order: T = (P0/R)/Rho (grid.cpp:882), cloud-in-cell
// at the particle (particle.cpp:402-456), RHS_T = (Tg - T_p) Nu pi Dp k
// (particle.cpp:460-474); also emits the deposit record of the particle
__global__ __launch_bounds__(kBlock) void k_temp_gather(TempArgs a) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  long p = a.perm[q];
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
  double v = (Tg - a.Tp[p]) * a.iTth;
  a.rhs_t[p] = v;
  a.rec[q] = make_double4(xl, yl, zl, v);
  a.rkey[q] = (ii & ix.mx) | ((jj & ix.my) << 10) | ((kk & ix.mz) << 20);
}


// ---- row-streaming pulls --------------------------------------------------
// One block per x-row (j,k) of target cells.  The block walks the 9 source
// rows (j+dj, k+dk); for each it stages the row's CSR and then its records,
// in chunks of kChunk, in LDS with coalesced loads, and every thread scans
// the source cells i-1, i, i+1 of its own cells from LDS.  Sources are
// visited in a fixed order (rows, then increasing position), so the sums are
// the same on every run and for any chunking.
constexpr int kChunk = 1024;
constexpr int kMaxNx = 512;

__device__ __forceinline__ void row_of_block(const Geo &g, int &j, int &k, unsigned &b) {
  b = xcd_block(blockIdx.x, gridDim.x);
  j = (int)(b % g.ny);
  k = (int)(b / g.ny);
}

// heat deposit (P9/P10): Q(c) = sum of w_c(p) RHS_T(p) over the records whose
// rounded cell r has c in {r-1, r} along every axis
__global__ __launch_bounds__(kBlock) void k_heat_rows(Geo g, const int32_t *__restrict__ start,
                                                      const double4 *__restrict__ rec,
                                                      const int32_t *__restrict__ rkey, double *out,
                                                      double *partials) {
  __shared__ int32_t sst[kMaxNx + 1];
  __shared__ double4 srec[kChunk];
  __shared__ int32_t skey[kChunk];
  __shared__ double red[kBlock];
  int j, k;
  unsigned b;
  row_of_block(g, j, k, b);
  Ix ix = make_ix(g);
  const int nx = g.nx;
  double acc[kMaxNx / kBlock] = {};
  for (int a = 0; a < 9; ++a) {
    int dk = a / 3 - 1, dj = a % 3 - 1;
    long R = ((long)((k + dk) & ix.mz) * g.ny + ((j + dj) & ix.my)) * nx;
    for (int t = threadIdx.x; t <= nx; t += blockDim.x) sst[t] = start[R + t];
    __syncthreads();
    int32_t rb = sst[0], re = sst[nx];
    for (int32_t c0 = rb; c0 < re; c0 += kChunk) {
      int cnt = re - c0 < kChunk ? re - c0 : kChunk;
      for (int t = threadIdx.x; t < cnt; t += blockDim.x) { srec[t] = rec[c0 + t]; skey[t] = rkey[c0 + t]; }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < kMaxNx / kBlock; ++u) {
        int i = threadIdx.x + u * kBlock;
        if (i >= nx) continue;
        int oi[3];
        sorted_offsets(i, nx, oi);
        for (int e = 0; e < 3; ++e) {
          int f = (i + oi[e]) & ix.mx;
          int32_t q0 = sst[f] > c0 ? sst[f] : c0, q1 = sst[f + 1] < c0 + cnt ? sst[f + 1] : c0 + cnt;
          for (int32_t q = q0; q < q1; ++q) {
            int32_t rk = skey[q - c0];
            int ox = (i - (rk & 1023) + 1) & ix.mx;  // 1: c is rx, 0: c is rx - 1
            int oy = (j - ((rk >> 10) & 1023) + 1) & ix.my;
            int oz = (k - ((rk >> 20) & 1023) + 1) & ix.mz;
            if (ox > 1 || oy > 1 || oz > 1) continue;
            double4 r4 = srec[q - c0];
            double wx = ox ? r4.x : (1 - r4.x), wy = oy ? r4.y : (1 - r4.y), wz = oz ? r4.z : (1 - r4.z);
            acc[u] += wx * wy * wz * r4.w;
          }
        }
      }
      __syncthreads();
    }
  }
  double mine = 0.0;
#pragma unroll
  for (int u = 0; u < kMaxNx / kBlock; ++u) {
    int i = threadIdx.x + u * kBlock;
    if (i < nx) {
      out[((long)k * g.ny + j) * nx + i] = acc[u];
      mine += acc[u];
    }
  }
  red[threadIdx.x] = mine;
  __syncthreads();
  for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[b] = red[0];
}

// velocity deposit (P6): S_a(c) = [sum over the particles of cell c of their
// (1 - frac_a) RHS_a] + [sum over cell c - e_a of frac_a RHS_a], each part in
// increasing sorted position.  Component 2a of contrib holds frac_a RHS_a,
// component 2a+1 holds (1 - frac_a) RHS_a.
__global__ __launch_bounds__(kBlock) void k_vel_rows(Geo g, const int32_t *__restrict__ start,
                                                     const double *__restrict__ contrib, long cs,
                                                     double *Sx, double *Sy, double *Sz) {
  __shared__ int32_t sst[kMaxNx + 1];
  __shared__ double sc[4][kChunk];
  int j, k;
  unsigned b;
  row_of_block(g, j, k, b);
  Ix ix = make_ix(g);
  const int nx = g.nx;
  constexpr int U = kMaxNx / kBlock;
  double own[U][3] = {}, nb[U][3] = {};
  // own row: components 0 (from cell i-1), 1, 3, 5 (from cell i)
  // row j-1: component 2 (from cell i); row k-1: component 4 (from cell i)
  for (int ph = 0; ph < 3; ++ph) {
    long R = ph == 0 ? ((long)k * g.ny + j) * nx
           : ph == 1 ? ((long)k * g.ny + ((j - 1) & ix.my)) * nx
                     : ((long)((k - 1) & ix.mz) * g.ny + j) * nx;
    for (int t = threadIdx.x; t <= nx; t += blockDim.x) sst[t] = start[R + t];
    __syncthreads();
    int32_t rb = sst[0], re = sst[nx];
    for (int32_t c0 = rb; c0 < re; c0 += kChunk) {
      int cnt = re - c0 < kChunk ? re - c0 : kChunk;
      for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
        if (ph == 0) {
          sc[0][t] = contrib[0 * cs + c0 + t];
          sc[1][t] = contrib[1 * cs + c0 + t];
          sc[2][t] = contrib[3 * cs + c0 + t];
          sc[3][t] = contrib[5 * cs + c0 + t];
        } else {
          sc[0][t] = contrib[(ph == 1 ? 2 : 4) * cs + c0 + t];
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < U; ++u) {
        int i = threadIdx.x + u * kBlock;
        if (i >= nx) continue;
        int32_t q0 = sst[i] > c0 ? sst[i] : c0, q1 = sst[i + 1] < c0 + cnt ? sst[i + 1] : c0 + cnt;
        if (ph == 0) {
          for (int32_t q = q0; q < q1; ++q) {
            own[u][0] += sc[1][q - c0];
            own[u][1] += sc[2][q - c0];
            own[u][2] += sc[3][q - c0];
          }
          int f = (i - 1) & ix.mx;
          int32_t r0 = sst[f] > c0 ? sst[f] : c0, r1 = sst[f + 1] < c0 + cnt ? sst[f + 1] : c0 + cnt;
          for (int32_t q = r0; q < r1; ++q) nb[u][0] += sc[0][q - c0];
        } else {
          for (int32_t q = q0; q < q1; ++q) nb[u][ph] += sc[0][q - c0];
        }
      }
      __syncthreads();
    }
  }
  long rowc = ((long)k * g.ny + j) * nx;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    int i = threadIdx.x + u * kBlock;
    if (i < nx) {
      Sx[rowc + i] = own[u][0] + nb[u][0];
      Sy[rowc + i] = own[u][1] + nb[u][1];
      Sz[rowc + i] = own[u][2] + nb[u][2];
    }
  }
}

// incremental binning: count (PLACE = false) or place (PLACE = true) the
// particles of the 27 base cells around each cell whose new key is that cell.
// Placement visits the base cells in increasing index and each cell's
// particles in increasing physical index, so the bins are stable.
template <bool PLACE>
__global__ __launch_bounds__(kBlock) void k_bin_rows(Geo g, const int32_t *__restrict__ start_base,
                                                     const int32_t *__restrict__ key, int32_t *cnt,
                                                     const int32_t *__restrict__ start_new, int32_t *perm,
                                                     int32_t *inv, const int32_t *moved) {
  if (moved[0]) return;
  __shared__ int32_t sst[kMaxNx + 1];
  __shared__ int32_t skey[kChunk];
  int j, k;
  unsigned b;
  row_of_block(g, j, k, b);
  Ix ix = make_ix(g);
  const int nx = g.nx;
  int32_t m[kMaxNx / kBlock] = {};
  int32_t w[kMaxNx / kBlock] = {};
  long rowc = ((long)k * g.ny + j) * nx;
  if (PLACE) {
#pragma unroll
    for (int u = 0; u < kMaxNx / kBlock; ++u) {
      int i = threadIdx.x + u * kBlock;
      if (i < nx) w[u] = start_new[rowc + i];
    }
  }
  int oj[3], ok[3];
  sorted_offsets(j, g.ny, oj);
  sorted_offsets(k, g.nz, ok);
  for (int a = 0; a < 9; ++a) {
    long R = ((long)((k + ok[a / 3]) & ix.mz) * g.ny + ((j + oj[a % 3]) & ix.my)) * nx;
    for (int t = threadIdx.x; t <= nx; t += blockDim.x) sst[t] = start_base[R + t];
    __syncthreads();
    int32_t rb = sst[0], re = sst[nx];
    for (int32_t c0 = rb; c0 < re; c0 += kChunk) {
      int cntc = re - c0 < kChunk ? re - c0 : kChunk;
      for (int t = threadIdx.x; t < cntc; t += blockDim.x) skey[t] = key[c0 + t];
      __syncthreads();
#pragma unroll
      for (int u = 0; u < kMaxNx / kBlock; ++u) {
        int i = threadIdx.x + u * kBlock;
        if (i >= nx) continue;
        int32_t me = (int32_t)(rowc + i);
        int oi[3];
        sorted_offsets(i, nx, oi);
        for (int e = 0; e < 3; ++e) {
          int f = (i + oi[e]) & ix.mx;
          int32_t q0 = sst[f] > c0 ? sst[f] : c0, q1 = sst[f + 1] < c0 + cntc ? sst[f + 1] : c0 + cntc;
          for (int32_t q = q0; q < q1; ++q) {
            if (skey[q - c0] != me) continue;
            if (PLACE) { perm[w[u]] = q; inv[q] = w[u]; ++w[u]; }
            else ++m[u];
          }
        }
      }
      __syncthreads();
    }
  }
  if (!PLACE) {
#pragma unroll
    for (int u = 0; u < kMaxNx / kBlock; ++u) {
      int i = threadIdx.x + u * kBlock;
      if (i < nx) cnt[rowc + i] = m[u];
    }
    if (b == 0 && threadIdx.x == 0) cnt[g.n] = 0;
  }
}

struct CicDepArgs {
  Geo g;
  const double *x, *y, *z;
  const int32_t *perm;
  const int32_t *start;
  double *out;
  double *partials;
};

// P5 pull (statistics only): concentration at x_int, weight one per particle
// (particle.cpp:522-555)
__global__ __launch_bounds__(kBlock) void k_concentration(CicDepArgs a) {
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
            int ox = (i - ii + 1) & ix.mx, oy = (j - jj + 1) & ix.my, oz = (k - kk + 1) & ix.mz;
            if (ox > 1 || oy > 1 || oz > 1) continue;
            double xl = x * g.idx - ii + 0.5, yl = y * g.idy - jj + 0.5, zl = z * g.idz - kk + 0.5;
            double wx = ox ? xl : (1 - xl), wy = oy ? yl : (1 - yl), wz = oz ? zl : (1 - zl);
            acc += wx * wy * wz;
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
  const int32_t *key;        // cell of each particle in the order being applied
  const double *src[7];
  double *dst_a[7];
  double *dst_b[7];
  const double *s1[3];
  double *d1[3];
  const int32_t *flag_s;
  int32_t *flag_d;
  int32_t *base_key;
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
  a.base_key[q] = a.key[s];
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
  for (int l = (int)levels.size() - 2; l >= 0; --l) {
    long len_l = levels[l].second;
    long nb = (len_l + kScanChunk - 1) / kScanChunk;
    k_scan_add<<<(unsigned)nb, kBlock, 0, c->stream>>>(levels[l].first, len_l, levels[l + 1].first);
    TPX_LAUNCH_CHECK();
  }
}

// Stable bins of the particles by key (one int per particle): perm (sorted
// position -> particle), inv (particle -> sorted position), start (CSR).
// Uses the incremental pass when the physical order is sorted by base cell
// and no particle left the 27 cells around its base cell; otherwise the
// general atomic counting sort.  Both paths are enqueued; the device flag
// `moved` selects which one does the work.
static void bin_particles(tpx_ctx *c, const int32_t *key, int32_t *perm, int32_t *inv, int32_t *start,
                          bool incremental_possible) {
  Particles &pt = c->pt;
  long n = pt.n, nc = c->g.n;
  unsigned gp = grid1d(n, kBlock), gc = grid1d(nc, kBlock), gc1 = grid1d(nc + 1, kBlock);
  if (!incremental_possible) k_set_i<<<1, 64, 0, c->stream>>>(pt.moved, 1, 1);
  // counts
  unsigned rows = (unsigned)((long)c->g.ny * c->g.nz);
  if (incremental_possible)
    k_bin_rows<false><<<rows, kBlock, 0, c->stream>>>(c->g, pt.start_base, key, pt.cursor, nullptr, nullptr,
                                                      nullptr, pt.moved);
  k_zero_gated<<<gc1, kBlock, 0, c->stream>>>(pt.cursor, nc + 1, pt.moved);
  k_count<<<gp, kBlock, 0, c->stream>>>(key, pt.cursor, n, pt.moved);
  TPX_LAUNCH_CHECK();
  launch_exclusive_scan(c, pt.cursor, nc + 1);
  k_copy_i<<<gc1, kBlock, 0, c->stream>>>(start, pt.cursor, nc + 1);
  // placement
  if (incremental_possible)
    k_bin_rows<true><<<rows, kBlock, 0, c->stream>>>(c->g, pt.start_base, key, nullptr, start, perm, inv,
                                                     pt.moved);
  k_place<<<gp, kBlock, 0, c->stream>>>(key, pt.cursor, perm, n, pt.moved);
  k_fixup<<<gc, kBlock, 0, c->stream>>>(start, perm, inv, nc, pt.moved);
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
  a.contrib = c->p.twc ? pt.contrib : nullptr;
  a.cstride = pt.cap;
  a.inv_int = pt.inv_int;
  a.base_key = pt.base_sorted ? pt.base_key : nullptr;
  a.moved = pt.moved;
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
  k_set_i<<<1, 64, 0, c->stream>>>(pt.moved, 0, 1);
  k_advance<<<grid1d(pt.n, kBlock), kBlock, 0, c->stream>>>(a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_PARTICLE);
  pt.np1_is_base = false;  // the np1 arrays now hold x_np1 = x + dt*post*u_int
}

void launch_velocity_deposit(tpx_ctx *c) {
  Particles &pt = c->pt;
  unsigned rows = (unsigned)((long)c->g.ny * c->g.nz);
  timing_begin(c, TPX_K_DEPOSIT);
  k_vel_rows<<<rows, kBlock, 0, c->stream>>>(c->g, pt.start_int, pt.contrib, pt.cap, c->S[0], c->S[1], c->S[2]);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_DEPOSIT);
}

// bin by the floor cell of x_new (keys written by the advance kernel)
void launch_sort_new(tpx_ctx *c) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  timing_begin(c, TPX_K_SORT);
  bin_particles(c, pt.key, pt.perm_new, pt.inv_new, pt.start_new, pt.base_sorted);
  timing_end(c, TPX_K_SORT);
}

// bin by the floor cell of one stage's positions with the general sort
void launch_sort_stage(tpx_ctx *c, int stage, bool into_int) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  int s = stage == TPX_BASE ? pt.st_base : stage == TPX_INT ? pt.st_int : stage == TPX_NEW ? pt.st_new : pt.st_np1;
  timing_begin(c, TPX_K_SORT);
  k_keys<<<grid1d(pt.n, kBlock), kBlock, 0, c->stream>>>(c->g, pt.q[0][s], pt.q[1][s], pt.q[2][s], pt.key, pt.n);
  TPX_LAUNCH_CHECK();
  bin_particles(c, pt.key, into_int ? pt.perm_int : pt.perm_new, into_int ? pt.inv_int : pt.inv_new,
                into_int ? pt.start_int : pt.start_new, false);
  timing_end(c, TPX_K_SORT);
}

// temperature gather + heat records, in the sorted order of `stage`
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
  a.perm = stage == TPX_INT ? pt.perm_int : pt.perm_new;
  a.rhs_t = pt.rhs_t;
  a.rec = pt.rec;
  a.rkey = pt.rkey;
  a.n = pt.n;
  timing_begin(c, TPX_K_PARTICLE);
  k_temp_gather<<<grid1d(pt.n, kBlock), kBlock, 0, c->stream>>>(a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_PARTICLE);
}

// heat deposit from the records of the last temperature gather
void launch_temp_deposit(tpx_ctx *c, int stage) {
  Particles &pt = c->pt;
  const int32_t *start = stage == TPX_INT ? pt.start_int : pt.start_new;
  timing_begin(c, TPX_K_DEPOSIT);
  if (pt.n > 0) {
    unsigned rows = (unsigned)((long)c->g.ny * c->g.nz);
    k_heat_rows<<<rows, kBlock, 0, c->stream>>>(c->g, start, pt.rec, pt.rkey, c->Q, c->partials);
    TPX_LAUNCH_CHECK();
    launch_finish_sum(c, (int)rows, 0, SC_SUM_Q);
  }
  timing_end(c, TPX_K_DEPOSIT);
}

void launch_concentration(tpx_ctx *c) {
  Particles &pt = c->pt;
  int s = pt.st_int;
  CicDepArgs a{};
  a.g = c->g;
  a.x = pt.q[0][s]; a.y = pt.q[1][s]; a.z = pt.q[2][s];
  a.perm = pt.perm_int;
  a.start = pt.start_int;
  a.out = c->C;
  a.partials = c->partials;
  unsigned nb = grid1d(c->g.n, kBlock);
  timing_begin(c, TPX_K_DEPOSIT);
  k_concentration<<<nb, kBlock, 0, c->stream>>>(a);
  TPX_LAUNCH_CHECK();
  launch_finish_sum(c, (int)nb, 0, SC_SUM_C);
  timing_end(c, TPX_K_DEPOSIT);
}

// Physically reorder the base state into the cell order described by
// perm_int (the bins of the current base positions), duplicate it into the
// int stage (equal to the base at a step boundary), and make that order the
// base order of the next step: perm_int/inv_int become the identity and
// start_base the CSR of start_int.
void launch_particle_permute(tpx_ctx *c) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  int sb = pt.st_base;
  int free_a = pt.st_np1, free_b = pt.st_int;  // np1 is stale (np1_is_base), int == base
  PermArgs a{};
  a.perm = pt.perm_int;
  a.key = pt.key;
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
  a.base_key = pt.base_key;
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
  launch_iota(c, pt.inv_int, pt.n);
  k_copy_i<<<grid1d(c->g.n + 1, kBlock), kBlock, 0, c->stream>>>(pt.start_base, pt.start_int, c->g.n + 1);
  TPX_LAUNCH_CHECK();
  pt.base_sorted = true;
  timing_end(c, TPX_K_SORT);
  check(hipGetLastError());
}

}  // namespace tpx
