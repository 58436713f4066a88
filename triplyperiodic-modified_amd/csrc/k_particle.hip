// k_particle.hip -- Lagrangian particle path of the RK4 substep (gfx950).
//
// Particles are SoA arrays kept physically ordered by the cell of their base
// position (re-sorted once per step).  Every substep they are re-binned by
// the floor cell of x_new with a counting sort: the advance kernel counts
// each particle into its cell with an integer atomic (no return value, so
// no wave waits on it), an exclusive scan turns the counts into CSR offsets,
// a placement pass takes slots from per-cell cursors, and a per-cell sort
// puts each cell's particles in increasing physical index.
// Integer counts do not depend on the order the atomics took, and the final
// sort removes the one thing that does, so the bins are identical on every
// run.
//
// Deposits are pull-form segment reductions, never floating-point atomics:
//   velocity (particle.cpp:494-520): face i of cell c sums the (1 - x_l)
//     shares of the particles in cell c and the x_l shares of cell c - e_x;
//     the advance kernel writes the shares at the particles' sorted
//     positions, so a block marching along z reads each plane's shares
//     contiguously into LDS and takes the neighbour-cell sums through LDS, a
//     halo row and the previous plane;
//   heat (particle.cpp:620-654): the same march gathers T at each plane's
//     particles (T of three planes staged in LDS) and keeps the per-particle
//     record (cloud-in-cell fractions, RHS_T, rounded cell) in LDS.  The
//     rounded cell of a particle in floor cell B is B or B + 1 along each
//     axis, so its 8 targets lie in the 3x3x3 block around B.  Each thread
//     sums the 27 weighted moments of its own cell's particles, the x shifts
//     are added through LDS, the z shifts in a register window, and a last
//     pass adds the y shifts.
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
  int32_t *cnt;              // per-cell counts of x_new (null: no binning this substep)
  int32_t *key;
  long n;
  int int_is_base, np1_is_base;  // the int / np1 inputs are the base arrays
  double c_post, c_pre;     // dt*post, dt*pre
  double iTp, g3[3];        // 1/Tp, ParticleGravity*g
  double iS, Dp, I0;        // 1/(mp Cvp), particle diameter, lamp intensity
  double epsp;              // scalar epsilonp
  double L[3];
  int rk3;
  __device__ void to_global() {
    for (int q = 0; q < 3; ++q) ru[q] = gptr(ru[q]);
    rho = gptr(rho);
    for (int q = 0; q < 7; ++q) base[q] = gptr(base[q]);
    for (int q = 0; q < 7; ++q) in_int[q] = gptr(in_int[q]);
    for (int q = 0; q < 7; ++q) np1_in[q] = gptr(np1_in[q]);
    for (int q = 0; q < 7; ++q) np1[q] = gptr(np1[q]);
    for (int q = 0; q < 7; ++q) nw[q] = gptr(nw[q]);
    for (int q = 0; q < 3; ++q) base_w[q] = gptr(base_w[q]);
    rhs_t = gptr(rhs_t);
    tau = gptr(tau);
    eps = gptr(eps);
    contrib = gptr(contrib);
    inv_int = gptr(inv_int);
    cnt = gptr(cnt);
    key = gptr(key);
  }
};

// P1-P4, P6 shares and P7 wrap for one particle (particle.cpp:314-399,
// 477-492, 494-520, 58-104)
// cnt[k] += 1 for every active lane, with one atomic per run of equal keys
// among consecutive lanes (the particles come nearly in cell order, and
// atomics on one cache line are served one after another).  The active lanes
// must be a prefix of the wave.
__device__ __forceinline__ void count_runs(int32_t *cnt, int k) {
  const int lane = (int)(threadIdx.x & 63);
  const int kprev = __shfl_up(k, 1, 64);
  const bool head = lane == 0 || kprev != k;
  const unsigned long long heads = __ballot(head);
  const int nact = __popcll(__ballot(true));
  if (head) {
    const unsigned long long above = heads & ~((2ull << lane) - 1);
    const int end = min(above ? __ffsll((long long)above) - 1 : 64, nact);
    atomicAdd(&cnt[k], end - lane);
  }
}

struct k_advance {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const AdvArgs &a) {
  long p = (long)xcd_block(blockIdx.x, gridDim.x) * block_dim_x() + threadIdx.x;
  if (p >= a.n) return;
  const Geo &g = a.g;
  Ix ix = make_ix(g);
  // stage values: base (always x, y, z; the rest unless the last substep),
  // int (x, y, z, u, v, w) and np1, each loaded once when stages coincide
  double b[7], vi[6], v1[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) b[q] = (q < 3 || !a.rk3) ? a.base[q][p] : 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) vi[q] = a.int_is_base ? b[q] : a.in_int[q][p];
#pragma unroll
  for (int q = 0; q < 7; ++q) v1[q] = a.np1_is_base ? b[q] : a.np1_in[q][p];
  double xi = vi[0], yi = vi[1], zi = vi[2];
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
    double ui = vi[3 + d];
    n1[d] = v1[d] + a.c_post * ui;
    nv[d] = a.rk3 ? n1[d] : b[d] + a.c_pre * ui;
  }
  // update_velocity (particle.cpp:362-399) and the deposit shares
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double ui = vi[3 + d];
    double rhs = (gas[d] - ui) * itp;
    n1[3 + d] = v1[3 + d] + a.c_post * (rhs + a.g3[d]);
    nv[3 + d] = a.rk3 ? n1[3 + d] : b[3 + d] + a.c_pre * (rhs + a.g3[d]);
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
    n1[6] = v1[6] + a.c_post * a.iS * (rt + alpha);
    nv[6] = a.rk3 ? n1[6] : b[6] + a.c_pre * a.iS * (rt + alpha);
  }
  // periodic wrap of x_new, x and x_np1 (particle.cpp:58-104); x_int keeps
  // its value unless it is the base buffer itself (first substep)
  double bx[3] = {b[0], b[1], b[2]};
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
  // the new and np1 stages stream out (read again only in the next substep,
  // long after they have left the cache): non-temporal stores keep them from
  // displacing the gas fields the gathers reuse
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    __builtin_nontemporal_store(n1[q], &a.np1[q][p]);
    __builtin_nontemporal_store(nv[q], &a.nw[q][p]);
  }
  int ki = (int)(nv[0] * g.idx), kj = (int)(nv[1] * g.idy), kz = (int)(nv[2] * g.idz);
  long nk = ix(ki, kj, kz);
  a.key[p] = (int32_t)nk;
  if (a.cnt) count_runs(a.cnt, (int)nk);  // only lanes past n left early: the active lanes are a prefix
}
};

// floor-cell key of one stage of positions, counted into its cell
struct k_keys {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(const Geo &g, const double *x, const double *y, const double *z, int32_t *key,
                       int32_t *cnt, long n) {
  long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
  if (p >= n) return;
  Ix ix = make_ix(g);
  int32_t k = (int32_t)ix((int)(x[p] * g.idx), (int)(y[p] * g.idy), (int)(z[p] * g.idz));
  key[p] = k;
  count_runs(cnt, k);  // only lanes past n left early
}
};

struct k_set_i {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(int32_t *a, int32_t v, long n) {
  long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
  if (p < n) a[p] = v;
}
};

struct k_copy_i {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(int32_t *__restrict__ d, const int32_t *__restrict__ s, long n) {
  long p = (long)blockIdx.x * block_dim_x() + threadIdx.x;
  if (p < n) d[p] = s[p];
}
};

// placement: each particle takes the next free slot of its cell's segment
// (cursor starts as a copy of the CSR offsets).  The particles arrive nearly
// in cell order, so a wave's keys come in runs of equal cells on a few cache
// lines, and returning atomics on one line are served one after another: each
// run of equal keys in a wave reserves its slots with one atomic (taken by
// the run's first lane) and its lanes take consecutive slots.  The order of
// the slots within a cell is undone by k_fixup.
struct k_place {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const int32_t *__restrict__ key, int32_t *__restrict__ cursor,
                                                  int32_t *__restrict__ perm, long n) {
  const long p = (long)xcd_block(blockIdx.x, gridDim.x) * block_dim_x() + threadIdx.x;
  const bool ok = p < n;  // the valid lanes of a wave are a prefix of it
  const int k = ok ? key[p] : -1;
  const int lane = (int)(threadIdx.x & 63);
  const int kprev = __shfl_up(k, 1, 64);
  const bool head = ok && (lane == 0 || kprev != k);
  const unsigned long long heads = __ballot(head);
  const int nvalid = __popcll(__ballot(ok));
  // the head of this lane's run: the highest head at or below it (a valid
  // lane's run starts at or before it: lane 0 is a head, and a non-head has
  // the key of the valid lane before it)
  const unsigned long long below = heads & ((2ull << lane) - 1);
  const int hl = ok ? 63 - __clzll(below) : lane;
  const unsigned long long above = heads & ~((2ull << hl) - 1);
  const int end = min(above ? __ffsll((long long)above) - 1 : 64, nvalid);
  int base = 0;
  if (head) base = atomicAdd(&cursor[k], end - hl);
  base = __shfl(base, hl, 64);
  if (ok) perm[base + (lane - hl)] = (int32_t)p;
}
};

__device__ __forceinline__ void cswap(int32_t &a, int32_t &b) {
  int32_t lo = a < b ? a : b, hi = a < b ? b : a;
  a = lo;
  b = hi;
}

// per-cell sort of the placed indices: the result is the stable order (by
// physical index) whatever order the placement atomics took.  Segments of
// up to 8 are sorted in registers by a fixed compare-exchange network
// (padded with INT32_MAX); longer ones by insertion in memory.
struct k_fixup {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(const int32_t *__restrict__ start, int32_t *perm, int32_t *inv, long ncell) {
  long c = (long)xcd_block(blockIdx.x, gridDim.x) * block_dim_x() + threadIdx.x;
  if (c >= ncell) return;
  int32_t b = start[c], e = start[c + 1];
  const int cnt = e - b;
  if (cnt <= 8) {
    int32_t v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = q < cnt ? perm[b + q] : INT32_MAX;
    // Batcher odd-even merge sort of 8 (19 compare-exchanges)
    cswap(v[0], v[1]); cswap(v[2], v[3]); cswap(v[4], v[5]); cswap(v[6], v[7]);
    cswap(v[0], v[2]); cswap(v[1], v[3]); cswap(v[4], v[6]); cswap(v[5], v[7]);
    cswap(v[1], v[2]); cswap(v[5], v[6]);
    cswap(v[0], v[4]); cswap(v[1], v[5]); cswap(v[2], v[6]); cswap(v[3], v[7]);
    cswap(v[2], v[4]); cswap(v[3], v[5]);
    cswap(v[1], v[2]); cswap(v[3], v[4]); cswap(v[5], v[6]);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < cnt) {
        perm[b + q] = v[q];
        inv[v[q]] = b + q;
      }
    return;
  }
  for (int32_t q = b + 1; q < e; ++q) {
    int32_t v = perm[q];
    int32_t r = q - 1;
    while (r >= b && perm[r] > v) { perm[r + 1] = perm[r]; --r; }
    perm[r + 1] = v;
  }
  for (int32_t q = b; q < e; ++q) inv[perm[q]] = q;
}
};

struct k_iota {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(int32_t *p, long n) {
  long q = (long)blockIdx.x * block_dim_x() + threadIdx.x;
  if (q < n) p[q] = (int32_t)q;
}
};

// ---- exclusive scan of int32 (three phases, block size 256 x 8) ----
constexpr int kScanPer = 8;
constexpr int kScanChunk = kBlock * kScanPer;

struct k_scan_blocks {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(int32_t *data, long n, int32_t *block_sums) {
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
};

struct k_scan_add {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(int32_t *data, long n, const int32_t *offs) {
  long q = (long)blockIdx.x * kScanChunk + threadIdx.x;
  int32_t o = offs[blockIdx.x];
  for (int e = 0; e < kScanPer; ++e, q += kBlock)
    if (q < n) data[q] += o;
}
};

// ---- deposits ----
constexpr int kHeatMaxBlock = 512;

// ---- temperature gather + heat deposit by per-cell moments (P8-P10) --------
// One block covers `rows` whole x-rows and marches over `zc` output planes
// (plus one halo plane on each side).  For every particle of a base plane it
// computes the cloud-in-cell fractions, the gas temperature T = (P0/R)/Rho
// (grid.cpp:882) at the particle from three planes of T staged in LDS (one
// division per cell instead of 8 per particle; particle.cpp:402-456) and
// RHS_T = (Tg - T_p) Nu pi Dp k (particle.cpp:460-474), which it stores by
// particle.  Every thread then sums, over its cell's particles in sorted
// order, the 27 products w_x(a) w_y(b) w_z(c) RHS_T for the target offsets
// a, b, c in {-1, 0, 1}.  The x offsets are exchanged through LDS, the z
// offsets accumulate in a three-plane register window, and the result for
// output cell (i, j + b, z) from this row is stored as t3[b + 1].
struct HeatFusedArgs {
  Geo g;
  const int32_t *start;      // CSR of the stage's cells
  const int32_t *perm;       // sorted position -> particle
  const double *x, *y, *z, *Tp;
  const double *rho;
  const double *scal;
  int p0_slot;
  double R, iTth;
  double *rhs_t;             // by particle
  double *t3;                // 3 x n: contribution of row j to row j - 1, j, j + 1
  int rows, zc;
  __device__ void to_global() {
    start = gptr(start);
    perm = gptr(perm);
    x = gptr(x);
    y = gptr(y);
    z = gptr(z);
    Tp = gptr(Tp);
    rho = gptr(rho);
    scal = gptr(scal);
    rhs_t = gptr(rhs_t);
    t3 = gptr(t3);
  }
};

// The block works on whole x-rows of one base plane at a time; the cells of
// those rows are consecutive, so their particles fill one consecutive range
// of sorted positions.  That range is taken in chunks of kHeatChunk x
// blockDim particles: every thread gathers T for its particles of the chunk
// (one or two, loads issued together) and leaves a record (cloud-in-cell
// fractions, RHS_T, upper-cell bits) in LDS, then every thread sums the
// moments of its own cell's records.
constexpr int kHeatChunk = 2;

struct k_heat_fused {
  static constexpr int kBounds = kHeatMaxBlock;
  static __device__ __forceinline__ void run(const HeatFusedArgs &a) {
  // LDS: 9 x blockDim doubles shared by the records of a chunk and the x
  // exchange, then T: 3 planes x (rows + 2) x nx
  extern __shared__ double xs[];
  const Geo &g = a.g;
  Ix ix = make_ix(g);
  const int nx = g.nx, bd = (int)block_dim_x(), t = (int)threadIdx.x;
  const int i = t & ix.mx, r = t / nx;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int nyb = g.ny / a.rows;
  const int j0 = (int)(b % (unsigned)nyb) * a.rows, j = j0 + r;
  const int z0 = (int)(b / (unsigned)nyb) * a.zc;
  const int il = r * nx + ((i - 1) & ix.mx), ir = r * nx + ((i + 1) & ix.mx);
  double4 *rec = reinterpret_cast<double4 *>(xs);                        // kHeatChunk x blockDim
  int32_t *rbit = reinterpret_cast<int32_t *>(xs + 4 * kHeatChunk * bd);  // kHeatChunk x blockDim
  double *tt = xs + 9 * bd;
  const int TPL = (a.rows + 2) * nx;  // doubles per staged plane, at most 3 per thread
  const double at = a.scal[a.p0_slot] / a.R;
  // Rho of plane kz, rows j0-1 .. j0+rows, this thread's share
  auto load_rho = [&](int kz, double (&rr)[3]) {
    const long zoff = (long)(kz & ix.mz) * g.nxy;
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int e = t + u * bd;
      if (e < TPL) {
        const int row = e / nx, col = e - row * nx;
        rr[u] = a.rho[zoff + (long)((j0 + row - 1) & ix.my) * nx + col];
      }
    }
  };
  // T = (P0/R)/Rho into ring slot s
  auto store_T = [&](const double (&rr)[3], int s) {
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int e = t + u * bd;
      if (e < TPL) tt[s * TPL + e] = at / rr[u];
    }
  };
  auto slot_of = [&](int kz) { return (kz - z0 + 3) % 3; };  // kz >= z0 - 2
  auto cell0 = [&](int kz) { return ((long)(kz & ix.mz) * g.ny + j0) * nx; };  // first cell of the rows
  {
    double r0[3], r1[3], r2[3];
    load_rho(z0 - 2, r0);
    load_rho(z0 - 1, r1);
    load_rho(z0, r2);
    store_T(r0, slot_of(z0 - 2));
    store_T(r1, slot_of(z0 - 1));
    store_T(r2, slot_of(z0));
  }
  // this plane's sorted ranges and the particles of its first chunk are
  // loaded one plane ahead
  int32_t Qa, Qb, q0, q1;
  {
    const long cb = cell0(z0 - 1);
    Qa = a.start[cb]; Qb = a.start[cb + bd];
    q0 = a.start[cb + t]; q1 = a.start[cb + t + 1];
  }
  int32_t pp[kHeatChunk];
  double px[kHeatChunk], py[kHeatChunk], pz[kHeatChunk], pq[kHeatChunk];
  auto load_chunk = [&](int32_t c0, int32_t ce) {
#pragma unroll
    for (int e = 0; e < kHeatChunk; ++e) {
      const int32_t q = c0 + e * bd + t;
      pp[e] = q < ce ? a.perm[q] : 0;
    }
#pragma unroll
    for (int e = 0; e < kHeatChunk; ++e)
      if (c0 + e * bd + t < ce) {
        px[e] = a.x[pp[e]]; py[e] = a.y[pp[e]]; pz[e] = a.z[pp[e]]; pq[e] = a.Tp[pp[e]];
      }
  };
  load_chunk(Qa, min(Qa + kHeatChunk * bd, Qb));
  __syncthreads();
  double sa[3] = {0., 0., 0.}, sb[3] = {0., 0., 0.};
  for (int zz = -1; zz <= a.zc; ++zz) {
    const int kabs = z0 + zz;
    const int sm = slot_of(kabs - 1), s0 = slot_of(kabs), sp = slot_of(kabs + 1);
    // next plane's ranges, and Rho of the plane staged after this one
    int32_t nQa = 0, nQb = 0, nq0 = 0, nq1 = 0;
    if (zz < a.zc) {
      const long cb = cell0(kabs + 1);
      nQa = a.start[cb]; nQb = a.start[cb + bd];
      nq0 = a.start[cb + t]; nq1 = a.start[cb + t + 1];
    }
    double rr[3];
    if (zz < a.zc) load_rho(kabs + 2, rr);
    // gather of one particle into record slot e
    auto gather = [&](int32_t p, double x, double y, double z, double tp, int e) {
      const double fx = x * g.idx, fy = y * g.idy, fz = z * g.idz;
      const int ii = (int)(round(fx)), jj = (int)(round(fy)), kk = (int)(round(fz));
      const double xl = fx - ii + 0.5, yl = fy - jj + 0.5, zl = fz - kk + 0.5;
      // base (floor) cell of the particle: the cell it is binned in
      const int bi = (int)fx, bj = (int)fy, bk = (int)fz;
      const int ux = ii - bi, uy = jj - bj, uz = kk - bk;
      const int ci = (bi & ix.mx) + ux, crow = (bj & ix.my) - j0 + 1 + uy;
      const int s1 = uz ? sp : s0, s2 = uz ? s0 : sm;  // planes kk and kk - 1
      // gas2part_Temp (particle.cpp:402-456): the trilinear weights of T at
      // (ii - a, jj - b, kk - c), a, b, c in {0, 1}, taken as nested
      // interpolations along x, y, z (the same weights; rounding differs
      // from the eight-term sum at the last bit)
      const int ch = ci & ix.mx, cl = (ci - 1) & ix.mx;
      const double *r11 = tt + s1 * TPL + crow * nx, *r10 = r11 - nx;  // plane kk, rows jj, jj - 1
      const double *r01 = tt + s2 * TPL + crow * nx, *r00 = r01 - nx;  // plane kk - 1
      auto lx = [&](const double *row) { return fma(xl, row[ch] - row[cl], row[cl]); };
      const double t1 = fma(yl, lx(r11) - lx(r10), lx(r10));
      const double t0 = fma(yl, lx(r01) - lx(r00), lx(r00));
      const double Tg = fma(zl, t1 - t0, t0);
      const double v = (Tg - tp) * a.iTth;
      a.rhs_t[p] = v;
      rec[e] = make_double4(xl, yl, zl, v);
      rbit[e] = (ux & 1) | ((uy & 1) << 1) | ((uz & 1) << 2);
    };
    double m[3][3][3];
#pragma unroll
    for (int u = 0; u < 27; ++u) m[u / 9][(u / 3) % 3][u % 3] = 0.0;
    for (int32_t c0 = Qa; c0 < Qb; c0 += kHeatChunk * bd) {
      const int32_t ce = min(c0 + kHeatChunk * bd, Qb);
      if (c0 != Qa) load_chunk(c0, ce);  // planes holding more than one chunk
#pragma unroll
      for (int e = 0; e < kHeatChunk; ++e)
        if (c0 + e * bd + t < ce) gather(pp[e], px[e], py[e], pz[e], pq[e], e * bd + t);
      __syncthreads();
      // the records of cell (i, j, k) inside this chunk, in sorted order
      const int32_t qe = min(q1, ce);
      for (int32_t q = max(q0, c0); q < qe; ++q) {
        const double4 v = rec[q - c0];
        const int ub = rbit[q - c0];
        const bool ux = ub & 1, uy = ub & 2, uz = ub & 4;
        const double wx[3] = {ux ? 0.0 : 1 - v.x, ux ? 1 - v.x : v.x, ux ? v.x : 0.0};
        const double wy[3] = {uy ? 0.0 : 1 - v.y, uy ? 1 - v.y : v.y, uy ? v.y : 0.0};
        const double wz[3] = {uz ? 0.0 : 1 - v.z, uz ? 1 - v.z : v.z, uz ? v.z : 0.0};
        // (w_x w_y RHS_T) w_z with fused multiply-adds: the moments are
        // deposit sums, compared with a tolerance, not bit for bit
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
#pragma unroll
          for (int ay = 0; ay < 3; ++ay) {
            const double wxyv = wx[ax] * wy[ay] * v.w;
#pragma unroll
            for (int az = 0; az < 3; ++az) m[ax][ay][az] = fma(wxyv, wz[az], m[ax][ay][az]);
          }
      }
      __syncthreads();  // the records are read before the next chunk or the exchange
    }
    // the next plane's first chunk, while this plane's moments are exchanged
    if (zz < a.zc) load_chunk(nQa, min(nQa + kHeatChunk * bd, nQb));
    // x exchange in two rounds of 9: the right neighbour's offset -1 moments,
    // then the left neighbour's offset +1 moments
    double R[3][3];
#pragma unroll
    for (int s = 0; s < 9; ++s) xs[s * bd + t] = m[0][s / 3][s % 3];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 9; ++s) R[s / 3][s % 3] = xs[s * bd + ir];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 9; ++s) xs[s * bd + t] = m[2][s / 3][s % 3];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 9; ++s) R[s / 3][s % 3] = (xs[s * bd + il] + m[1][s / 3][s % 3]) + R[s / 3][s % 3];
    if (zz >= 1) {
      const long co = ((long)(z0 + zz - 1) * g.ny + j) * nx + i;
#pragma unroll
      for (int e = 0; e < 3; ++e) a.t3[e * g.n + co] = sa[e] + R[e][0];
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      sa[e] = sb[e] + R[e][1];
      sb[e] = R[e][2];
    }
    // plane k - 1 is no longer needed: stage plane k + 2 into its slot
    if (zz < a.zc) store_T(rr, sm);
    __syncthreads();
    Qa = nQa; Qb = nQb; q0 = nq0; q1 = nq1;
  }
}
};

// Stage 2: Q(i, j, k) = t3[2](i, j-1, k) + t3[1](i, j, k) + t3[0](i, j+1, k),
// with block partial sums for the mean of Q
struct k_heat_combine {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const Geo &g, const double *__restrict__ t3, double *out,
                                                         double *partials) {
  __shared__ double red[kBlock];
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const long c = (long)b * block_dim_x() + threadIdx.x;
  double v = 0.0;
  if (c < g.n) {
    const int my = g.ny - 1;
    const long row = c / g.nx;
    const int j = (int)(row % g.ny);
    const long base = c - (long)j * g.nx;
    v = (t3[2 * g.n + base + (long)((j - 1) & my) * g.nx] + t3[g.n + c]) + t3[base + (long)((j + 1) & my) * g.nx];
    out[c] = v;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = block_dim_x() / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[b] = red[0];
}
};

// velocity deposit (P6): S_a(c) = [sum over the particles of cell c of their
// (1 - frac_a) RHS_a] + [sum over cell c - e_a of frac_a RHS_a], each part in
// increasing sorted position.  Component 2a of contrib holds frac_a RHS_a,
// component 2a+1 holds (1 - frac_a) RHS_a.
// Sums of NV per-particle arrays over the particles of each of `ncell`
// consecutive cells starting at cell `cb`, one cell per thread (threads past
// `ncell` get zeros), each sum taken in sorted order from 0.  The cells'
// particles occupy one consecutive range of sorted positions; it is read in
// chunks of kVelChunk x blockDim positions, every thread loading its one or
// two positions of every array (coalesced) into LDS, then every thread adding
// up its own cell's positions from LDS.  All threads of the block call it.
constexpr int kVelChunk = 2;

template <int NV>
__device__ __forceinline__ void cell_sums(const int32_t *__restrict__ start, long cb, int ncell,
                                          const double *const (&arr)[NV], double *lds, double (&out)[NV]) {
  const int t = (int)threadIdx.x, bd = (int)block_dim_x(), cap = kVelChunk * bd;
  const int32_t Qa = start[cb], Qb = start[cb + ncell];
  int32_t q0 = 0, q1 = 0;
  if (t < ncell) { q0 = start[cb + t]; q1 = start[cb + t + 1]; }
#pragma unroll
  for (int v = 0; v < NV; ++v) out[v] = 0.0;
  for (int32_t c0 = Qa; c0 < Qb; c0 += cap) {
    const int32_t ce = min(c0 + cap, Qb);
    double val[kVelChunk][NV];
#pragma unroll
    for (int e = 0; e < kVelChunk; ++e) {
      const int32_t q = c0 + e * bd + t;
#pragma unroll
      for (int v = 0; v < NV; ++v) val[e][v] = q < ce ? arr[v][q] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < kVelChunk; ++e)
#pragma unroll
      for (int v = 0; v < NV; ++v) lds[v * cap + e * bd + t] = val[e][v];
    __syncthreads();
    const int32_t qe = min(q1, ce);
    for (int32_t q = max(q0, c0); q < qe; ++q)
#pragma unroll
      for (int v = 0; v < NV; ++v) out[v] += lds[v * cap + (q - c0)];
    __syncthreads();  // LDS is free again
  }
}

// P6 reduction by planes (particle.cpp:494-520): the block covers `rows`
// whole x-rows and `zc` planes.  For every cell the sums of the six share
// arrays over its own particles give the shares of its low faces (c1, c3,
// c5); the shares of the high faces of the cell below along x, y, z (c0, c2,
// c4 of cells (i-1, j, k), (i, j-1, k), (i, j, k-1)) come from the
// neighbouring thread through LDS, from the halo row j0 - 1, and from the
// previous plane of the march.
struct k_vel_planes {
  static constexpr int kBounds = kHeatMaxBlock;
  static __device__ __forceinline__ void run(const Geo &g, const int32_t *__restrict__ start,
                                                             const double *__restrict__ contrib, long cs,
                                                             double *Sx, double *Sy, double *Sz, int rows,
                                                             int zc) {
  extern __shared__ double ls[];  // 6 x kVelChunk x blockDim
  Ix ix = make_ix(g);
  const int nx = g.nx, bd = (int)block_dim_x(), t = (int)threadIdx.x;
  const int i = t & ix.mx, r = t / nx;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int nyb = g.ny / rows;
  const int j0 = (int)(b % (unsigned)nyb) * rows, j = j0 + r;
  const int z0 = (int)(b / (unsigned)nyb) * zc;
  const double *c[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) c[v] = contrib + v * cs;
  const double *const c4[1] = {c[4]};
  const double *const c2[1] = {c[2]};
  // c4 sums of plane z0 - 1
  double below[1];
  cell_sums<1>(start, ((long)((z0 - 1) & ix.mz) * g.ny + j0) * nx, bd, c4, ls, below);
  for (int zz = 0; zz < zc; ++zz) {
    const int k = z0 + zz;
    double own[6], halo[1];
    cell_sums<6>(start, ((long)k * g.ny + j0) * nx, bd, c, ls, own);
    cell_sums<1>(start, ((long)k * g.ny + ((j0 - 1) & ix.my)) * nx, nx, c2, ls, halo);
    ls[t] = own[0];
    ls[bd + t] = own[2];
    if (t < nx) ls[2 * bd + t] = halo[0];
    __syncthreads();
    const double nxs = ls[r * nx + ((i - 1) & ix.mx)];
    const double nys = r > 0 ? ls[bd + (r - 1) * nx + i] : ls[2 * bd + i];
    const long co = ((long)k * g.ny + j) * nx + i;
    Sx[co] = own[1] + nxs;
    Sy[co] = own[3] + nys;
    Sz[co] = own[5] + below[0];
    below[0] = own[4];
    __syncthreads();
  }
}
};

struct CicDepArgs {
  Geo g;
  const double *x, *y, *z;
  const int32_t *perm;
  const int32_t *start;
  double *out;
  double *partials;
  __device__ void to_global() {
    x = gptr(x);
    y = gptr(y);
    z = gptr(z);
    perm = gptr(perm);
    start = gptr(start);
    out = gptr(out);
    partials = gptr(partials);
  }
};

// P5 pull (statistics only): concentration at x_int, weight one per particle
// (particle.cpp:522-555)
struct k_concentration {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const CicDepArgs &a) {
  __shared__ double sh[kBlock];
  const Geo &g = a.g;
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  long c = (long)b * block_dim_x() + threadIdx.x;
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
  for (int s = block_dim_x() / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && a.partials) a.partials[b] = sh[0];
}
};

struct PermArgs {
  const int32_t *perm;       // sorted position -> particle
  const double *src[7];
  double *dst[7];
  const double *s1[3];
  double *d1[3];
  const int32_t *flag_s;
  int32_t *flag_d;
  long n;
  __device__ void to_global() {
    perm = gptr(perm);
    for (int q = 0; q < 7; ++q) src[q] = gptr(src[q]);
    for (int q = 0; q < 7; ++q) dst[q] = gptr(dst[q]);
    for (int q = 0; q < 3; ++q) s1[q] = gptr(s1[q]);
    for (int q = 0; q < 3; ++q) d1[q] = gptr(d1[q]);
    flag_s = gptr(flag_s);
    flag_d = gptr(flag_d);
  }
};

// physical re-sort into cell order: sorted position q gathers the base state
// of particle perm[q] (the writes stay contiguous; a scatter of the same data
// measured slower)
struct k_permute {
  static constexpr int kBounds = kBlock;
  static __device__ __forceinline__ void run(const PermArgs &a) {
  long q = (long)xcd_block(blockIdx.x, gridDim.x) * block_dim_x() + threadIdx.x;
  if (q >= a.n) return;
  long s = a.perm[q];
#pragma unroll
  for (int d = 0; d < 7; ++d) a.dst[d][q] = a.src[d][s];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (a.s1[d]) a.d1[d][q] = a.s1[d][s];
  a.flag_d[q] = a.flag_s[s];
}
};

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static void check(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

void launch_iota(tpx_ctx *c, int32_t *p, int64_t n) {
  if (n <= 0) return;
  tpx_launch<k_iota>(grid1d(n, kBlock), kBlock, 0, c->stream, p, n);
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
    tpx_launch<k_scan_blocks>((unsigned)nb, kBlock, 0, c->stream, cur, len, tmp);
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
    tpx_launch<k_scan_add>((unsigned)nb, kBlock, 0, c->stream, levels[l].first, len_l, levels[l + 1].first);
    TPX_LAUNCH_CHECK();
  }
}

// Stable bins of the particles by key (one int per particle): perm (sorted
// position -> particle), inv (particle -> sorted position), start (CSR).
// On entry `start` holds the per-cell counts (start[n_cell] = 0).
static void bin_particles(tpx_ctx *c, const int32_t *key, int32_t *perm, int32_t *inv, int32_t *start) {
  Particles &pt = c->pt;
  long n = pt.n, nc = c->g.n;
  launch_exclusive_scan(c, start, nc + 1);
  tpx_launch<k_copy_i>(grid1d(nc + 1, kBlock), kBlock, 0, c->stream, pt.cursor, start, nc + 1);
  tpx_launch<k_place>(grid1d(n, kBlock), kBlock, 0, c->stream, key, pt.cursor, perm, n);
  tpx_launch<k_fixup>(grid1d(nc, kBlock), kBlock, 0, c->stream, start, perm, inv, nc);
  TPX_LAUNCH_CHECK();
}

static void zero_counts(tpx_ctx *c, int32_t *cnt) {
  tpx_launch<k_set_i>(grid1d(c->g.n + 1, kBlock), kBlock, 0, c->stream, cnt, 0, c->g.n + 1);
  TPX_LAUNCH_CHECK();
}

void launch_bin_by_key(tpx_ctx *c, const int32_t *key, int32_t *perm, int32_t *inv, int32_t *start) {
  bin_particles(c, key, perm, inv, start);
}

void launch_zero_counts(tpx_ctx *c, int32_t *cnt) { zero_counts(c, cnt); }

void launch_particle_advance(tpx_ctx *c, const Coef &k, const double *const ru_int[3],
                             const double *rho_int, bool count_new) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  AdvArgs a{};
  a.g = c->g;
  for (int d = 0; d < 3; ++d) a.ru[d] = ru_int[d];
  a.rho = rho_int;
  int sb = pt.st_base, si = pt.int_buf();
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
  a.cnt = count_new ? pt.start_new : nullptr;
  a.key = pt.key;
  a.n = pt.n;
  a.int_is_base = pt.int_is_base || si == sb;
  a.np1_is_base = pt.np1_is_base;
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
  if (count_new) {
    timing_begin(c, TPX_K_SORT);
    zero_counts(c, pt.start_new);
    timing_end(c, TPX_K_SORT);
  }
  timing_begin(c, TPX_K_ADVANCE);
  tpx_launch<k_advance>(grid1d(pt.n, kBlock), kBlock, 0, c->stream, a);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_ADVANCE);
  pt.np1_is_base = false;  // the np1 arrays now hold x_np1 = x + dt*post*u_int
  pt.new_counted = count_new;
}

// Launch shape of the heat moment and velocity deposit kernels: `rows` whole
// x-rows per block and `zc` planes per z-march.  The built-in choice is 256
// threads (nx when nx > 256) and up to 32 planes (8 for the velocity
// deposit); a tuned or set shape replaces it.
struct HeatShape { int threads, rows, zc; unsigned blocks; };
static HeatShape deposit_shape(const tpx_ctx *c, int kind) {
  const Geo &g = c->g;
  const TileTune::Shape &set = kind == 0 ? c->tile.heat : c->tile.vel;
  HeatShape h;
  h.rows = set.rows > 0 ? set.rows : (g.nx >= 256 ? 1 : std::min(256 / g.nx, g.ny));
  h.zc = set.zc > 0 ? set.zc : std::min(kind == 0 ? 32 : 8, g.nz);
  h.threads = g.nx * h.rows;
  h.blocks = (unsigned)((g.ny / h.rows) * (g.nz / h.zc));
  return h;
}

static size_t deposit_lds(const Geo &g, int kind, int rows) {
  const size_t threads = (size_t)g.nx * rows;
  return kind == 0 ? sizeof(double) * (9 * threads + 3 * (size_t)(rows + 2) * g.nx)
                   : sizeof(double) * 6 * kVelChunk * threads;
}

bool tile_shape_valid(const tpx_ctx *c, int kind, int rows, int zc) {
  const Geo &g = c->g;
  return rows >= 1 && zc >= 1 && g.nx * rows <= kHeatMaxBlock && g.ny % rows == 0 && g.nz % zc == 0 &&
         deposit_lds(g, kind, rows) <= c->max_lds;
}

void tile_tune_begin(tpx_ctx *c) {
  TileTune &t = c->tile;
  for (auto &p : t.pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  t.pending.clear();
  for (int k = 0; k < 2; ++k) {
    t.cand[k].clear();
    for (int rows = 1; rows <= 8; rows *= 2)
      for (int zc = 2; zc <= 32; zc *= 2)
        if (tile_shape_valid(c, k, rows, zc)) t.cand[k].push_back({rows, zc});
    t.best_ms[k].assign(t.cand[k].size(), -1.0);
    t.runs[k].assign(t.cand[k].size(), 0);
    t.next[k] = 0;
  }
  t.active = true;
}

// the shape for this launch: the next candidate while tuning, with its events
static void tile_pick(tpx_ctx *c, int kind, hipEvent_t *ev) {
  TileTune &t = c->tile;
  ev[0] = ev[1] = nullptr;
  if (!t.active || g_rec || t.cand[kind].empty()) return;
  const size_t i = t.next[kind]++ % t.cand[kind].size();
  (kind == 0 ? t.heat : t.vel) = t.cand[kind][i];
  if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess)
    throw std::runtime_error("hipEventCreate");
  t.pending.push_back({kind, i, ev[0], ev[1]});
}

void tile_tune_collect(tpx_ctx *c) {
  TileTune &t = c->tile;
  if (!t.active) return;
  if (hipStreamSynchronize(c->stream) != hipSuccess) throw std::runtime_error("hipStreamSynchronize");
  for (auto &p : t.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      double &b = t.best_ms[p.kind][p.cand];
      if (b < 0 || ms < b) b = ms;
      ++t.runs[p.kind][p.cand];
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  t.pending.clear();
  // done once each kernel that ran has timed all its candidates kRuns times
  // (the fastest run of each counts); a kernel that did not run keeps the
  // built-in shape
  bool done = true;
  for (int k = 0; k < 2; ++k) {
    if (t.next[k] == 0) continue;
    for (int r : t.runs[k]) done = done && r >= TileTune::kRuns;
  }
  const bool any = t.next[0] > 0 || t.next[1] > 0;
  if (!done || !any) return;
  for (int k = 0; k < 2; ++k) {
    TileTune::Shape &s = k == 0 ? t.heat : t.vel;
    if (t.next[k] == 0) {
      s = {0, 0};
      continue;
    }
    size_t best = 0;
    for (size_t i = 1; i < t.cand[k].size(); ++i)
      if (t.best_ms[k][i] < t.best_ms[k][best]) best = i;
    s = t.cand[k][best];
  }
  t.active = false;
}

void launch_velocity_deposit(tpx_ctx *c) {
  Particles &pt = c->pt;
  timing_begin(c, TPX_K_DEPOSIT);
  const Geo &g = c->g;
  hipEvent_t ev[2];
  tile_pick(c, 1, ev);
  HeatShape h = deposit_shape(c, 1);
  if (!tile_shape_valid(c, 1, h.rows, h.zc))
    throw std::logic_error("velocity deposit launch shape does not tile the grid or fit in LDS");
  const size_t lds = deposit_lds(g, 1, h.rows);
  if (ev[0]) check(hipEventRecord(ev[0], c->stream));
  tpx_launch<k_vel_planes>(h.blocks, h.threads, lds, c->stream, g, pt.start_int, pt.contrib, pt.cap, c->S[0], c->S[1],
                                                        c->S[2], h.rows, h.zc);
  TPX_LAUNCH_CHECK();
  if (ev[1]) check(hipEventRecord(ev[1], c->stream));
  timing_end(c, TPX_K_DEPOSIT);
}

// bin by the floor cell of x_new (keys and counts written by the advance kernel)
void launch_sort_new(tpx_ctx *c) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  if (!pt.new_counted) throw std::logic_error("x_new was not counted by the particle advance");
  timing_begin(c, TPX_K_SORT);
  bin_particles(c, pt.key, pt.perm_new, pt.inv_new, pt.start_new);
  timing_end(c, TPX_K_SORT);
  pt.new_counted = false;
}

// bin by the floor cell of one stage's positions
void launch_sort_stage(tpx_ctx *c, int stage, bool into_int) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  int s = stage == TPX_BASE ? pt.st_base : stage == TPX_INT ? pt.int_buf() : stage == TPX_NEW ? pt.st_new : pt.st_np1;
  int32_t *start = into_int ? pt.start_int : pt.start_new;
  timing_begin(c, TPX_K_SORT);
  zero_counts(c, start);
  tpx_launch<k_keys>(grid1d(pt.n, kBlock), kBlock, 0, c->stream, c->g, pt.q[0][s], pt.q[1][s], pt.q[2][s], pt.key, start,
                                                         pt.n);
  TPX_LAUNCH_CHECK();
  bin_particles(c, pt.key, into_int ? pt.perm_int : pt.perm_new, into_int ? pt.inv_int : pt.inv_new, start);
  timing_end(c, TPX_K_SORT);
}

// temperature gather + heat records, in the sorted order of `stage`
// temperature gather + heat deposit in one pass over the stage's bins
void launch_heat_exchange(tpx_ctx *c, int stage, const double *rho, int p0_slot) {
  Particles &pt = c->pt;
  timing_begin(c, TPX_K_DEPOSIT);
  if (pt.n > 0) {
    const Geo &g = c->g;
    hipEvent_t ev[2];
    tile_pick(c, 0, ev);
    HeatShape h = deposit_shape(c, 0);
    if (!tile_shape_valid(c, 0, h.rows, h.zc))
      throw std::logic_error("heat deposit launch shape does not tile the grid or fit in LDS");
    const int s = stage == TPX_INT ? pt.int_buf() : pt.st_new;
    HeatFusedArgs a{};
    a.g = g;
    a.start = stage == TPX_INT ? pt.start_int : pt.start_new;
    a.perm = stage == TPX_INT ? pt.perm_int : pt.perm_new;
    a.x = pt.q[0][s]; a.y = pt.q[1][s]; a.z = pt.q[2][s]; a.Tp = pt.q[6][s];
    a.rho = rho;
    a.scal = c->scal;
    a.p0_slot = p0_slot;
    a.R = c->R;
    a.iTth = c->p.nu * TPX_PI * c->p.dp * c->p.k;
    a.rhs_t = pt.rhs_t;
    a.t3 = c->dep_tmp;
    a.rows = h.rows;
    a.zc = h.zc;
    const size_t lds = deposit_lds(g, 0, h.rows);
    if (ev[0]) check(hipEventRecord(ev[0], c->stream));
    tpx_launch<k_heat_fused>(h.blocks, h.threads, lds, c->stream, a);
    TPX_LAUNCH_CHECK();
    if (ev[1]) check(hipEventRecord(ev[1], c->stream));
    unsigned nb = grid1d(g.n, kBlock);
    tpx_launch<k_heat_combine>(nb, kBlock, 0, c->stream, g, c->dep_tmp, c->Q, c->partials);
    TPX_LAUNCH_CHECK();
    launch_finish_sum(c, (int)nb, 0, SC_SUM_Q);
  }
  timing_end(c, TPX_K_DEPOSIT);
}

void launch_concentration(tpx_ctx *c) {
  Particles &pt = c->pt;
  int s = pt.int_buf();
  CicDepArgs a{};
  a.g = c->g;
  a.x = pt.q[0][s]; a.y = pt.q[1][s]; a.z = pt.q[2][s];
  a.perm = pt.perm_int;
  a.start = pt.start_int;
  a.out = c->C;
  a.partials = c->partials;
  unsigned nb = grid1d(c->g.n, kBlock);
  timing_begin(c, TPX_K_DEPOSIT);
  tpx_launch<k_concentration>(nb, kBlock, 0, c->stream, a);
  TPX_LAUNCH_CHECK();
  launch_finish_sum(c, (int)nb, 0, SC_SUM_C);
  timing_end(c, TPX_K_DEPOSIT);
}

// Physically reorder the base state into the cell order described by
// perm_int (the bins of the current base positions), duplicate it into the
// int stage (equal to the base at a step boundary), and make that order the
// base order of the next step: perm_int/inv_int become the identity and
// start_int stays the CSR of that order.
void launch_particle_permute(tpx_ctx *c) {
  Particles &pt = c->pt;
  if (pt.n == 0) return;
  int sb = pt.st_base;
  int dst = pt.st_np1;  // np1 is stale (np1_is_base)
  PermArgs a{};
  a.perm = pt.perm_int;
  for (int q = 0; q < 7; ++q) {
    a.src[q] = pt.q[q][sb];
    a.dst[q] = pt.q[q][dst];
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
  tpx_launch<k_permute>(grid1d(pt.n, kBlock), kBlock, 0, c->stream, a);
  TPX_LAUNCH_CHECK();
  for (int d = 0; d < 3; ++d)
    if (use[d]) std::swap(*own[d], pt.spare[d]);
  std::swap(pt.flag, pt.itmp);
  pt.st_base = dst;
  pt.st_np1 = sb;
  pt.np1_is_base = true;
  // the int stage equals the base stage: read it from the base buffer and
  // keep st_int as the spare the first TimeAdvance_RK4 gives to the new stage
  if (pt.st_int == dst || pt.st_int == sb) {
    int used[3] = {pt.st_base, pt.st_np1, pt.st_new};
    for (int b = 0; b < 4; ++b)
      if (b != used[0] && b != used[1] && b != used[2]) pt.st_int = b;
  }
  pt.int_is_base = true;
  launch_iota(c, pt.perm_int, pt.n);
  launch_iota(c, pt.inv_int, pt.n);
  timing_end(c, TPX_K_SORT);
  check(hipGetLastError());
}

}  // namespace tpx
