// k_fft.hip -- the FFT-preconditioned Poisson iteration (poisson.cpp:391-462)
// as five batched LDS radix-2 FFT passes per iteration (gfx950).
//
// One iteration of the reference is
//   r  = (div_F2C(C grad_C2F P) - RHS) * cc,   C = 1/Rho_face, cc = -1/I_F2C(C)
//   dp = -Re IFFT( FFT(r) / lambda(k) ) / N^3   (zero mode passed through)
//   P += dp,  err = max|dp|
// with r real.  The passes here are
//   X  (forward): residual computed in registers, real-to-half-complex FFT
//                 along x; only kx = 0..nx/2 is stored ((nx/2+1) ny nz).
//   Y  (forward): complex FFT along y on tiles of 8 kx columns.
//   Z  (both):    forward FFT along z, divide by lambda, inverse FFT along z.
//   Y  (inverse).
//   X  (inverse): Hermitian extension, inverse FFT along x, P += dp and the
//                 block maxima of |dp|.
// C and cc are recomputed from Rho_new in the X pass instead of being stored.
#include <math.h>

#include <stdexcept>
#include <vector>

#include "tpx_internal.h"

namespace tpx {

namespace {

constexpr int kThreads = 256;
constexpr int kTile = 8;  // kx columns per block in the y/z passes

struct Ix {
  int mx, my, mz, nx;
  long nxy;
  __device__ __forceinline__ long operator()(int i, int j, int k) const {
    return (long)(k & mz) * nxy + (long)(j & my) * nx + (i & mx);
  }
};

__device__ __forceinline__ int ilog2(int n) { return 31 - __clz(n); }
__device__ __forceinline__ int bitrev(int v, int lg) { return (int)(__brev((unsigned)v) >> (32 - lg)); }

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-place radix-2 DIT transforms of `lines` lines of length n held in LDS
// (line l at buf + l*n, already in bit-reversed order).  tw[m] = e^{-2 pi i m/n}
// for m < n/2; INV uses the conjugates.
template <bool INV>
__device__ void lds_fft(double2 *buf, int n, int lines, const double2 *__restrict__ tw) {
  int half_n = n >> 1;
  int total = half_n * lines;
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, step = n / len;
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
      int l = b / half_n, bf = b - l * half_n;
      int grp = bf / half, pos = bf - grp * half;
      int i0 = grp * len + pos, i1 = i0 + half;
      double2 w = tw[pos * step];
      if (INV) w.y = -w.y;
      double2 *a = buf + (long)l * n;
      double2 u = a[i0], v = cmul(a[i1], w);
      a[i0] = make_double2(u.x + v.x, u.y + v.y);
      a[i1] = make_double2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
  }
}

struct XArgs {
  Geo g;
  const double *P;
  const double *rho;     // Rho_new
  const double *rhs;     // raw Poisson RHS
  const double *scal;    // SC_SUM_RHS
  double2 *spec;
  const double2 *tw;     // x twiddles
  double ncell;
};

// C on face A of cell (i,j,k): 1./Rho_face (grid.cpp:853)
template <int A>
__device__ __forceinline__ double Cface(const Ix &ix, const double *__restrict__ rho, int i, int j, int k) {
  long p = ix(i, j, k), q = ix(i - (A == 0), j - (A == 1), k - (A == 2));
  return 1. / ((rho[p] + rho[q]) * 0.5);
}

// residual r = (div_F2C(C grad_C2F P) - RHS) * cc at cell (i,j,k)
__device__ __forceinline__ double residual(const XArgs &a, const Ix &ix, int i, int j, int k, double mean) {
  const double *__restrict__ rho = a.rho;
  const double *__restrict__ P = a.P;
  const Geo &g = a.g;
  double cx0 = Cface<0>(ix, rho, i, j, k), cx1 = Cface<0>(ix, rho, i + 1, j, k);
  double cy0 = Cface<1>(ix, rho, i, j, k), cy1 = Cface<1>(ix, rho, i, j + 1, k);
  double cz0 = Cface<2>(ix, rho, i, j, k), cz1 = Cface<2>(ix, rho, i, j, k + 1);
  double third = 1. / 3.;
  double cc = -1. / (third * ((cx0 + cx1) * 0.5 + (cy0 + cy1) * 0.5 + (cz0 + cz1) * 0.5));
  double pc = P[ix(i, j, k)];
  double gx0 = ((pc - P[ix(i - 1, j, k)]) * g.idx) * cx0, gx1 = ((P[ix(i + 1, j, k)] - pc) * g.idx) * cx1;
  double gy0 = ((pc - P[ix(i, j - 1, k)]) * g.idy) * cy0, gy1 = ((P[ix(i, j + 1, k)] - pc) * g.idy) * cy1;
  double gz0 = ((pc - P[ix(i, j, k - 1)]) * g.idz) * cz0, gz1 = ((P[ix(i, j, k + 1)] - pc) * g.idz) * cz1;
  double rn = (gx1 - gx0) * g.idx + (gy1 - gy0) * g.idy + (gz1 - gz0) * g.idz;
  double RHS = a.rhs[ix(i, j, k)] - mean;  // RHS_Pois after make_mean_zero
  rn = rn - RHS;
  return rn * cc;
}

// X forward pass: one x line per group of threads
__global__ __launch_bounds__(kThreads) void k_xfwd(XArgs a, int tpl, int lpb) {
  extern __shared__ double2 lds[];
  const Geo &g = a.g;
  int nx = g.nx, lg = ilog2(nx), mxh = nx / 2 + 1;
  long nlines = (long)g.ny * g.nz;
  Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
  double mean = a.scal[SC_SUM_RHS] / a.ncell;
  int lt = threadIdx.x / tpl, t = threadIdx.x - lt * tpl;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  long line = (long)blk * lpb + lt;
  double2 *buf = lds + (long)lt * nx;
  if (line < nlines) {
    int j = (int)(line % g.ny), k = (int)(line / g.ny);
    for (int i = t; i < nx; i += tpl) buf[bitrev(i, lg)] = make_double2(residual(a, ix, i, j, k, mean), 0.0);
  }
  __syncthreads();
  lds_fft<false>(lds, nx, lpb, a.tw);
  if (line < nlines) {
    double2 *out = a.spec + line * mxh;
    for (int m = t; m < mxh; m += tpl) out[m] = buf[m];
  }
}

struct XInvArgs {
  Geo g;
  double *P;
  const double2 *spec;
  const double2 *tw;
  double *partials;
  double ncell;
};

// X inverse pass: Hermitian extension, inverse FFT, P += dp, block max|dp|
__global__ __launch_bounds__(kThreads) void k_xinv(XInvArgs a, int tpl, int lpb) {
  extern __shared__ double2 lds[];
  __shared__ double red[kThreads];
  const Geo &g = a.g;
  int nx = g.nx, lg = ilog2(nx), mxh = nx / 2 + 1;
  long nlines = (long)g.ny * g.nz;
  int lt = threadIdx.x / tpl, t = threadIdx.x - lt * tpl;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  long line = (long)blk * lpb + lt;
  double2 *buf = lds + (long)lt * nx;
  if (line < nlines) {
    const double2 *in = a.spec + line * mxh;
    for (int m = t; m < mxh; m += tpl) {
      double2 v = in[m];
      if (m == 0 || m == nx / 2) v.y = 0.0;
      buf[bitrev(m, lg)] = v;
      if (m > 0 && m < nx / 2) buf[bitrev(nx - m, lg)] = make_double2(v.x, -v.y);
    }
  }
  __syncthreads();
  lds_fft<true>(lds, nx, lpb, a.tw);
  double mx = 0.0;
  if (line < nlines) {
    double *Pl = a.P + line * nx;
    for (int i = t; i < nx; i += tpl) {
      double dp = (-1) * buf[i].x / (a.ncell);
      Pl[i] = Pl[i] + dp;
      mx = fmax(mx, fabs(dp));
    }
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) a.partials[blk] = red[0];
}

// Y pass (forward or inverse) on a tile of kTile kx columns at fixed z
template <bool INV>
__global__ __launch_bounds__(kThreads) void k_ypass(double2 *spec, int mxh, int ny, int nz,
                                                    const double2 *__restrict__ tw) {
  extern __shared__ double2 lds[];
  int ntile = (mxh + kTile - 1) / kTile;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  int tile = blk % ntile, k = blk / ntile;
  int m0 = tile * kTile;
  int lg = ilog2(ny);
  int cols = min(kTile, mxh - m0);
  double2 *base = spec + (long)k * ny * mxh + m0;
  for (int e = threadIdx.x; e < ny * kTile; e += blockDim.x) {
    int j = e / kTile, col = e - j * kTile;
    double2 v = col < cols ? base[(long)j * mxh + col] : make_double2(0.0, 0.0);
    lds[col * ny + bitrev(j, lg)] = v;
  }
  __syncthreads();
  lds_fft<INV>(lds, ny, kTile, tw);
  for (int e = threadIdx.x; e < ny * kTile; e += blockDim.x) {
    int j = e / kTile, col = e - j * kTile;
    if (col < cols) base[(long)j * mxh + col] = lds[col * ny + j];
  }
}

// Z pass: forward FFT along z, divide by lambda (poisson.cpp:433-451), inverse
__global__ __launch_bounds__(kThreads) void k_zpass(double2 *spec, int mxh, int ny, int nz,
                                                    const double2 *__restrict__ tw,
                                                    const double *__restrict__ lamx,
                                                    const double *__restrict__ lamy,
                                                    const double *__restrict__ lamz) {
  extern __shared__ double2 lds[];
  int ntile = (mxh + kTile - 1) / kTile;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  int tile = blk % ntile, j = blk / ntile;
  int m0 = tile * kTile;
  int lg = ilog2(nz);
  int cols = min(kTile, mxh - m0);
  long stride = (long)ny * mxh;
  double2 *base = spec + (long)j * mxh + m0;
  for (int e = threadIdx.x; e < nz * kTile; e += blockDim.x) {
    int k = e / kTile, col = e - k * kTile;
    double2 v = col < cols ? base[(long)k * stride + col] : make_double2(0.0, 0.0);
    lds[col * nz + bitrev(k, lg)] = v;
  }
  __syncthreads();
  lds_fft<false>(lds, nz, kTile, tw);
  for (int e = threadIdx.x; e < nz * kTile; e += blockDim.x) {
    int col = e / nz, k = e - col * nz;
    int m = m0 + col;
    if (col < cols && (m + j + k > 0)) {
      double lam = lamx[m] + lamy[j] + lamz[k];
      double2 v = lds[col * nz + k];
      v.y /= lam;
      v.x /= lam;
      lds[col * nz + k] = v;
    }
  }
  __syncthreads();
  // bit-reverse each line again for the inverse DIT transform
  for (int e = threadIdx.x; e < nz * kTile; e += blockDim.x) {
    int col = e / nz, k = e - col * nz;
    int r = bitrev(k, lg);
    if (r > k) {
      double2 t0 = lds[col * nz + k];
      lds[col * nz + k] = lds[col * nz + r];
      lds[col * nz + r] = t0;
    }
  }
  __syncthreads();
  lds_fft<true>(lds, nz, kTile, tw);
  for (int e = threadIdx.x; e < nz * kTile; e += blockDim.x) {
    int k = e / kTile, col = e - k * kTile;
    if (col < cols) base[(long)k * stride + col] = lds[col * nz + k];
  }
}

__global__ __launch_bounds__(1024) void k_finish_max(const double *__restrict__ partials, long nparts,
                                                     double *scal, int slot) {
  __shared__ double sh[1024];
  long per = (nparts + 1023) / 1024;
  long b = (long)threadIdx.x * per, e = b + per < nparts ? b + per : nparts;
  double v = 0.0;
  for (long q = b; q < e; ++q) v = fmax(v, partials[q]);
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[slot] = sh[0];
}

void line_shape(int n, int &tpl, int &lpb) {
  tpl = n / 2 < kThreads ? n / 2 : kThreads;
  if (tpl < 1) tpl = 1;
  lpb = kThreads / tpl;
}

}  // namespace

int fft_setup(tpx_ctx *c) {
  const Geo &g = c->g;
  int n[3] = {g.nx, g.ny, g.nz};
  double L[3] = {c->p.lx, c->p.ly, c->p.lz};
  double d[3] = {g.dx, g.dy, g.dz};
  for (int a = 0; a < 3; ++a) {
    std::vector<double2> tw(n[a] / 2 > 0 ? n[a] / 2 : 1);
    for (int m = 0; m < n[a] / 2; ++m) {
      double ang = 2.0 * 3.141592653589793 * (double)m / (double)n[a];
      tw[m] = make_double2(cos(ang), -sin(ang));
    }
    // lambda tables (poisson.cpp:433-449); the x table covers kx = 0..nx/2
    int cnt = a == 0 ? n[0] / 2 + 1 : n[a];
    std::vector<double> lam(cnt);
    double two_pi_over_l = (2 * 3.141592653589793) / L[a];
    double d2 = 2. / (d[a] * d[a]);
    for (int i = 0; i < cnt; ++i) {
      double w = ((i < n[a] / 2) ? i : i - n[a]) * two_pi_over_l;
      lam[i] = (d2) * (1 - cos(w * d[a]));
    }
    if (hipMalloc(&c->tw[a], sizeof(double2) * tw.size()) != hipSuccess) return TPX_E_OOM;
    if (hipMalloc(&c->lam[a], sizeof(double) * lam.size()) != hipSuccess) return TPX_E_OOM;
    if (hipMemcpy(c->tw[a], tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice) != hipSuccess) return TPX_E_HIP;
    if (hipMemcpy(c->lam[a], lam.data(), sizeof(double) * lam.size(), hipMemcpyHostToDevice) != hipSuccess) return TPX_E_HIP;
  }
  long mxh = g.nx / 2 + 1;
  if (hipMalloc(&c->spec[0], sizeof(double2) * mxh * g.ny * g.nz) != hipSuccess) return TPX_E_OOM;
  return TPX_OK;
}

void launch_poisson_iteration(tpx_ctx *c, const double *rho_new) {
  const Geo &g = c->g;
  int mxh = g.nx / 2 + 1;
  long nlines = (long)g.ny * g.nz;
  int tpl, lpb;
  line_shape(g.nx, tpl, lpb);
  unsigned nbx = (unsigned)((nlines + lpb - 1) / lpb);
  size_t shx = sizeof(double2) * (size_t)g.nx * lpb;
  timing_begin(c, TPX_K_FFT);
  XArgs xa{g, c->P, rho_new, c->rhs, c->scal, c->spec[0], c->tw[0], (double)g.n};
  k_xfwd<<<nbx, kThreads, shx, c->stream>>>(xa, tpl, lpb);
  TPX_LAUNCH_CHECK();
  int ntile = (mxh + kTile - 1) / kTile;
  unsigned nby = (unsigned)(ntile * g.nz), nbz = (unsigned)(ntile * g.ny);
  k_ypass<false><<<nby, kThreads, sizeof(double2) * kTile * g.ny, c->stream>>>(c->spec[0], mxh, g.ny, g.nz, c->tw[1]);
  TPX_LAUNCH_CHECK();
  k_zpass<<<nbz, kThreads, sizeof(double2) * kTile * g.nz, c->stream>>>(c->spec[0], mxh, g.ny, g.nz, c->tw[2],
                                                                               c->lam[0], c->lam[1], c->lam[2]);
  TPX_LAUNCH_CHECK();
  k_ypass<true><<<nby, kThreads, sizeof(double2) * kTile * g.ny, c->stream>>>(c->spec[0], mxh, g.ny, g.nz, c->tw[1]);
  TPX_LAUNCH_CHECK();
  XInvArgs xi{g, c->P, c->spec[0], c->tw[0], c->partials, (double)g.n};
  k_xinv<<<nbx, kThreads, shx, c->stream>>>(xi, tpl, lpb);
  TPX_LAUNCH_CHECK();
  k_finish_max<<<1, 1024, 0, c->stream>>>(c->partials, nbx, c->scal, SC_MAX_DP);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_FFT);
}

}  // namespace tpx
