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

struct Ix {
  int mx, my, mz, nx;
  long nxy;
  __device__ __forceinline__ long operator()(int i, int j, int k) const {
    return (long)(k & mz) * nxy + (long)(j & my) * nx + (i & mx);
  }
};

__device__ __forceinline__ int ilog2(int n) { return 31 - __clz(n); }

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Stockham autosort FFT of `lines` lines of length n held in LDS, natural
// order in and out: one radix-2 pass when log2(n) is odd, then radix-4
// passes.  Line-major layout (ILV false): element j of line l at l*n + j;
// interleaved layout (ILV true): at j*lines + l, with consecutive threads on
// consecutive lines, so both layouts are read and written without LDS bank
// conflicts.  Each pass reads its butterfly inputs into registers,
// waits, and writes the outputs in place.  tw[t] = e^{-2 pi i t/n}, t < n;
// INV uses the conjugates.  At most kMaxBf butterflies per thread per pass.
constexpr int kMaxBf = 4;

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

template <bool INV, bool ILV>
__device__ void lds_fft(double2 *buf, int n, int lines, const double2 *__restrict__ tw) {
  // element j of line l
  auto at = [&](int l, int j) -> double2 & { return ILV ? buf[j * lines + l] : buf[l * n + j]; };
  int lg = ilog2(n);
  int Ns = 1;
  if (lg & 1) {  // radix-2 pass with Ns = 1 (no twiddles)
    int per = n >> 1, total = per * lines;
    double2 y0[kMaxBf], y1[kMaxBf];
    int dst[kMaxBf];
#pragma unroll
    for (int q = 0; q < kMaxBf; ++q) {
      int b = threadIdx.x + q * blockDim.x;
      if (b < total) {
        int l = ILV ? b % lines : b / per, j = ILV ? b / lines : b - (b / per) * per;
        double2 a0 = at(l, j), a1 = at(l, j + per);
        y0[q] = cadd(a0, a1);
        y1[q] = csub(a0, a1);
        dst[q] = l | ((2 * j) << 16);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxBf; ++q) {
      int b = threadIdx.x + q * blockDim.x;
      if (b < total) {
        int l = dst[q] & 0xffff, d = dst[q] >> 16;
        at(l, d) = y0[q];
        at(l, d + 1) = y1[q];
      }
    }
    __syncthreads();
    Ns = 2;
  }
  int per = n >> 2, total = per * lines;
  for (; Ns < n; Ns <<= 2) {
    double2 y[kMaxBf][4];
    int dst[kMaxBf];
    int tstep = n / (4 * Ns);
#pragma unroll
    for (int q = 0; q < kMaxBf; ++q) {
      int b = threadIdx.x + q * blockDim.x;
      if (b < total) {
        int l = ILV ? b % lines : b / per, j = ILV ? b / lines : b - (b / per) * per;
        int k = j & (Ns - 1);
        double2 a0 = at(l, j), a1 = at(l, j + per), a2 = at(l, j + 2 * per), a3 = at(l, j + 3 * per);
        if (Ns > 1) {
          double2 w1 = tw[k * tstep], w2 = tw[2 * k * tstep], w3 = tw[3 * k * tstep];
          if (INV) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
          a1 = cmul(a1, w1); a2 = cmul(a2, w2); a3 = cmul(a3, w3);
        }
        double2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
        y[q][0] = cadd(t0, t2);
        y[q][2] = csub(t0, t2);
        if (!INV) {  // multiply t3 by -i
          y[q][1] = make_double2(t1.x + t3.y, t1.y - t3.x);
          y[q][3] = make_double2(t1.x - t3.y, t1.y + t3.x);
        } else {     // by +i
          y[q][1] = make_double2(t1.x - t3.y, t1.y + t3.x);
          y[q][3] = make_double2(t1.x + t3.y, t1.y - t3.x);
        }
        dst[q] = l | (((j - k) * 4 + k) << 16);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxBf; ++q) {
      int b = threadIdx.x + q * blockDim.x;
      if (b < total) {
        int l = dst[q] & 0xffff, d = dst[q] >> 16;
        at(l, d) = y[q][0];
        at(l, d + Ns) = y[q][1];
        at(l, d + 2 * Ns) = y[q][2];
        at(l, d + 3 * Ns) = y[q][3];
      }
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

// X forward pass: lpb consecutive x lines per block
__global__ __launch_bounds__(kThreads) void k_xfwd(XArgs a, int lpb) {
  extern __shared__ double2 lds[];
  const Geo &g = a.g;
  int nx = g.nx, mxh = nx / 2 + 1;
  long nlines = (long)g.ny * g.nz;
  Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
  double mean = a.scal[SC_SUM_RHS] / a.ncell;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  long line0 = (long)blk * lpb;
  for (int e = threadIdx.x; e < lpb * nx; e += blockDim.x) {
    int l = e / nx, i = e - l * nx;
    long line = line0 + l;
    double r = 0.0;
    if (line < nlines) r = residual(a, ix, i, (int)(line % g.ny), (int)(line / g.ny), mean);
    lds[e] = make_double2(r, 0.0);
  }
  __syncthreads();
  lds_fft<false, false>(lds, nx, lpb, a.tw);
  for (int e = threadIdx.x; e < lpb * mxh; e += blockDim.x) {
    int l = e / mxh, m = e - l * mxh;
    long line = line0 + l;
    if (line < nlines) a.spec[line * mxh + m] = lds[l * nx + m];
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
__global__ __launch_bounds__(kThreads) void k_xinv(XInvArgs a, int lpb) {
  extern __shared__ double2 lds[];
  __shared__ double red[kThreads];
  const Geo &g = a.g;
  int nx = g.nx, mxh = nx / 2 + 1;
  long nlines = (long)g.ny * g.nz;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  long line0 = (long)blk * lpb;
  for (int e = threadIdx.x; e < lpb * mxh; e += blockDim.x) {
    int l = e / mxh, m = e - l * mxh;
    long line = line0 + l;
    double2 v = line < nlines ? a.spec[line * mxh + m] : make_double2(0.0, 0.0);
    if (m == 0 || m == nx / 2) v.y = 0.0;
    lds[l * nx + m] = v;
    if (m > 0 && m < nx / 2) lds[l * nx + nx - m] = make_double2(v.x, -v.y);
  }
  __syncthreads();
  lds_fft<true, false>(lds, nx, lpb, a.tw);
  double mx = 0.0;
  for (int e = threadIdx.x; e < lpb * nx; e += blockDim.x) {
    int l = e / nx, i = e - l * nx;
    long line = line0 + l;
    if (line < nlines) {
      double dp = (-1) * lds[e].x / (a.ncell);
      double *Pp = a.P + line * nx + i;
      *Pp = *Pp + dp;
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
                                                    const double2 *__restrict__ tw, int kTile) {
  extern __shared__ double2 lds[];
  int ntile = (mxh + kTile - 1) / kTile;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  int tile = blk % ntile, k = blk / ntile;
  int m0 = tile * kTile;
  int cols = min(kTile, mxh - m0);
  double2 *base = spec + (long)k * ny * mxh + m0;
  for (int e = threadIdx.x; e < ny * kTile; e += blockDim.x) {
    int j = e / kTile, col = e - j * kTile;
    double2 v = col < cols ? base[(long)j * mxh + col] : make_double2(0.0, 0.0);
    lds[e] = v;  // element j of column col at j*kTile + col
  }
  __syncthreads();
  lds_fft<INV, true>(lds, ny, kTile, tw);
  for (int e = threadIdx.x; e < ny * kTile; e += blockDim.x) {
    int j = e / kTile, col = e - j * kTile;
    if (col < cols) base[(long)j * mxh + col] = lds[e];
  }
}

// Z pass: forward FFT along z, divide by lambda (poisson.cpp:433-451), inverse
__global__ __launch_bounds__(kThreads) void k_zpass(double2 *spec, int mxh, int ny, int nz,
                                                    const double2 *__restrict__ tw,
                                                    const double *__restrict__ lamx,
                                                    const double *__restrict__ lamy,
                                                    const double *__restrict__ lamz, int kTile) {
  extern __shared__ double2 lds[];
  int ntile = (mxh + kTile - 1) / kTile;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  int tile = blk % ntile, j = blk / ntile;
  int m0 = tile * kTile;
  int cols = min(kTile, mxh - m0);
  long stride = (long)ny * mxh;
  double2 *base = spec + (long)j * mxh + m0;
  for (int e = threadIdx.x; e < nz * kTile; e += blockDim.x) {
    int k = e / kTile, col = e - k * kTile;
    double2 v = col < cols ? base[(long)k * stride + col] : make_double2(0.0, 0.0);
    lds[e] = v;  // element k of column col at k*kTile + col
  }
  __syncthreads();
  lds_fft<false, true>(lds, nz, kTile, tw);
  for (int e = threadIdx.x; e < nz * kTile; e += blockDim.x) {
    int k = e / kTile, col = e - k * kTile;
    int m = m0 + col;
    if (col < cols && (m + j + k > 0)) {
      double lam = lamx[m] + lamy[j] + lamz[k];
      double2 v = lds[e];
      v.y /= lam;
      v.x /= lam;
      lds[e] = v;
    }
  }
  __syncthreads();
  lds_fft<true, true>(lds, nz, kTile, tw);
  for (int e = threadIdx.x; e < nz * kTile; e += blockDim.x) {
    int k = e / kTile, col = e - k * kTile;
    if (col < cols) base[(long)k * stride + col] = lds[e];
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

// lines per block of the x passes and kx columns per block of the y/z
// passes: 2048 points per block keeps every FFT pass within kMaxBf
// butterflies per thread and the LDS tile at 32 KB
int lines_per_block(int n) {
  int l = 2048 / n;
  return l > 8 ? 8 : (l < 1 ? 1 : l);
}

}  // namespace

int fft_setup(tpx_ctx *c) {
  const Geo &g = c->g;
  int n[3] = {g.nx, g.ny, g.nz};
  double L[3] = {c->p.lx, c->p.ly, c->p.lz};
  double d[3] = {g.dx, g.dy, g.dz};
  for (int a = 0; a < 3; ++a) {
    std::vector<double2> tw(n[a]);
    for (int m = 0; m < n[a]; ++m) {
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
  int lpb = lines_per_block(g.nx);
  unsigned nbx = (unsigned)((nlines + lpb - 1) / lpb);
  size_t shx = sizeof(double2) * (size_t)g.nx * lpb;
  timing_begin(c, TPX_K_FFT);
  XArgs xa{g, c->P, rho_new, c->rhs, c->scal, c->spec[0], c->tw[0], (double)g.n};
  k_xfwd<<<nbx, kThreads, shx, c->stream>>>(xa, lpb);
  TPX_LAUNCH_CHECK();
  int ty = lines_per_block(g.ny), tz = lines_per_block(g.nz);
  unsigned nby = (unsigned)(((mxh + ty - 1) / ty) * g.nz), nbz = (unsigned)(((mxh + tz - 1) / tz) * g.ny);
  size_t shy = sizeof(double2) * (size_t)ty * g.ny, shz = sizeof(double2) * (size_t)tz * g.nz;
  k_ypass<false><<<nby, kThreads, shy, c->stream>>>(c->spec[0], mxh, g.ny, g.nz, c->tw[1], ty);
  TPX_LAUNCH_CHECK();
  k_zpass<<<nbz, kThreads, shz, c->stream>>>(c->spec[0], mxh, g.ny, g.nz, c->tw[2], c->lam[0], c->lam[1],
                                              c->lam[2], tz);
  TPX_LAUNCH_CHECK();
  k_ypass<true><<<nby, kThreads, shy, c->stream>>>(c->spec[0], mxh, g.ny, g.nz, c->tw[1], ty);
  TPX_LAUNCH_CHECK();
  XInvArgs xi{g, c->P, c->spec[0], c->tw[0], c->partials, (double)g.n};
  k_xinv<<<nbx, kThreads, shx, c->stream>>>(xi, lpb);
  TPX_LAUNCH_CHECK();
  k_finish_max<<<1, 1024, 0, c->stream>>>(c->partials, nbx, c->scal, SC_MAX_DP);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_FFT);
}

}  // namespace tpx
