// k_fft.hip -- the FFT-preconditioned Poisson iteration (poisson.cpp:391-462)
// as three (or five) batched FFT passes per iteration (gfx950).
//
// One iteration of the reference is
//   r  = (div_F2C(C grad_C2F P) - RHS) * cc,   C = 1/Rho_face, cc = -1/I_F2C(C)
//   dp = -Re IFFT( FFT(r) / lambda(k) ) / N^3   (zero mode passed through)
//   P += dp,  err = max|dp|
// with r real.  When every axis length is 16..256 the iteration is three
// passes over the half spectrum (k_xy_fwd, k_yz_pass, k_xy_inv below: the y
// transform is split between them along its four-step factorisation).
// Otherwise it is five:
//   X  (forward): residual computed in registers, real-to-half-complex FFT
//                 along x; only kx = 0..nx/2 is stored ((nx/2+1) ny nz).
//   Y  (forward): complex FFT along y on tiles of kx columns.
//   Z  (both):    forward FFT along z, divide by lambda, inverse FFT along z.
//   Y  (inverse).
//   X  (inverse): Hermitian extension, inverse FFT along x, P += dp and the
//                 block maxima of |dp|.
// Axis lengths 16..256 use four-step register kernels (n = R1 R2, one LDS
// exchange); 8 and 512 use LDS Stockham kernels.  C and cc are recomputed
// from Rho_new in the X pass instead of being stored.
#include <math.h>

#include <algorithm>
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
      int b = threadIdx.x + q * block_dim_x();
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
      int b = threadIdx.x + q * block_dim_x();
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
      int b = threadIdx.x + q * block_dim_x();
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
      int b = threadIdx.x + q * block_dim_x();
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
  const double *done;    // nonzero: the solve has converged, the iteration is skipped
  const double *res;     // the residual formed by k_residual_coef (a solve with the
                         // caller's face coefficient), or null: formed here from rho
  const double *in = nullptr;  // three-pass form: transform this field's rows instead of a
                               // residual (P, rho, rhs unread)
  __device__ void to_global() {
    P = gptr(P);
    rho = gptr(rho);
    rhs = gptr(rhs);
    scal = gptr(scal);
    spec = gptr(spec);
    tw = gptr(tw);
    done = gptr(done);
    res = gptr(res);
    in = gptr(in);
  }
};

// 1/x from the hardware reciprocal estimate and two Newton steps,
// r <- r + r (1 - x r): the relative error squares each step, so the result
// is within an ulp or so of the correctly rounded quotient, for a fraction
// of the instructions of an IEEE division.  Used only inside the Poisson
// iteration, whose result is compared with a tolerance (it is an FFT).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// C on face A of cell (i,j,k): 1./Rho_face (grid.cpp:853)
template <int A>
__device__ __forceinline__ double Cface(const Ix &ix, const double *__restrict__ rho, int i, int j, int k) {
  long p = ix(i, j, k), q = ix(i - (A == 0), j - (A == 1), k - (A == 2));
  return rcp_nr((rho[p] + rho[q]) * 0.5);
}

// residual r = (div_F2C(C grad_C2F P) - RHS) * cc from P at the cell and its
// six neighbours, C on the cell's six faces and the raw RHS
__device__ __forceinline__ double residual_of(const Geo &g, double pc, double pxm, double pxp, double pym,
                                              double pyp, double pzm, double pzp, double cx0, double cx1,
                                              double cy0, double cy1, double cz0, double cz1, double rhs,
                                              double mean) {
  double third = 1. / 3.;
  double cc = -rcp_nr(third * ((cx0 + cx1) * 0.5 + (cy0 + cy1) * 0.5 + (cz0 + cz1) * 0.5));
  double gx0 = ((pc - pxm) * g.idx) * cx0, gx1 = ((pxp - pc) * g.idx) * cx1;
  double gy0 = ((pc - pym) * g.idy) * cy0, gy1 = ((pyp - pc) * g.idy) * cy1;
  double gz0 = ((pc - pzm) * g.idz) * cz0, gz1 = ((pzp - pc) * g.idz) * cz1;
  double rn = (gx1 - gx0) * g.idx + (gy1 - gy0) * g.idy + (gz1 - gz0) * g.idz;
  double RHS = rhs - mean;  // RHS_Pois after make_mean_zero
  rn = rn - RHS;
  return rn * cc;
}

// the residual at cell (i,j,k)
__device__ __forceinline__ double residual(const XArgs &a, const Ix &ix, int i, int j, int k, double mean) {
  if (a.res) return a.res[ix(i, j, k)];
  const double *__restrict__ rho = a.rho;
  const double *__restrict__ P = a.P;
  return residual_of(a.g, P[ix(i, j, k)], P[ix(i - 1, j, k)], P[ix(i + 1, j, k)], P[ix(i, j - 1, k)],
                     P[ix(i, j + 1, k)], P[ix(i, j, k - 1)], P[ix(i, j, k + 1)], Cface<0>(ix, rho, i, j, k),
                     Cface<0>(ix, rho, i + 1, j, k), Cface<1>(ix, rho, i, j, k), Cface<1>(ix, rho, i, j + 1, k),
                     Cface<2>(ix, rho, i, j, k), Cface<2>(ix, rho, i, j, k + 1), a.rhs[ix(i, j, k)], mean);
}

// v moved along the 16-lane rows of the wave: CTRL 0x121 (row_ror:1) gives
// each lane the value of the lane before it in its row, 0x12F (row_ror:15)
// that of the lane after it (the first and last lanes wrap within the row)
template <int CTRL>
__device__ __forceinline__ double row_rotate(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The residuals of one whole x-row as the four-step x transform holds it:
// the 16 lanes ta of a wave's row take x = ta + 16 m, m < R2 (nx = 16 R2).
// The x neighbours of a point are then the lanes beside it (or, at the ends
// of the lane row, the lane at the other end one m over), so P and Rho at
// x - 1 and x + 1, and C on the face above x, come from the neighbouring
// lanes instead of memory: 11 loads and 6 reciprocals a point instead of 15
// and 7, the same arithmetic.  Every lane of the wave must be active.
template <int R2>
__device__ __forceinline__ void row_residuals(const XArgs &a, const Ix &ix, int ta, int j, int z, double mean,
                                              double2 (&v)[R2]) {
  const double *__restrict__ rho = a.rho;
  const double *__restrict__ P = a.P;
  const double *__restrict__ rhs = a.rhs;
  const long rc = (long)(z & ix.mz) * ix.nxy + (long)(j & ix.my) * ix.nx;
  const long ryl = (long)(z & ix.mz) * ix.nxy + (long)((j - 1) & ix.my) * ix.nx;
  const long ryh = (long)(z & ix.mz) * ix.nxy + (long)((j + 1) & ix.my) * ix.nx;
  const long rzl = (long)((z - 1) & ix.mz) * ix.nxy + (long)(j & ix.my) * ix.nx;
  const long rzh = (long)((z + 1) & ix.mz) * ix.nxy + (long)(j & ix.my) * ix.nx;
  // value at x - 1 / x + 1 of the point whose own value is cur, given the
  // value of the previous / next point of the lane
  auto lo_of = [&](double cur, double prev) {
    const double u = row_rotate<0x121>(cur), w = row_rotate<0x121>(prev);
    return ta == 0 ? w : u;
  };
  auto hi_of = [&](double cur, double next) {
    const double u = row_rotate<0x12F>(cur), w = row_rotate<0x12F>(next);
    return ta == 15 ? w : u;
  };
  double pp = P[rc + ta + 16 * (R2 - 1)], rp = rho[rc + ta + 16 * (R2 - 1)];  // the point before m = 0
  double pcur = P[rc + ta], rcur = rho[rc + ta];
  double cx0 = rcp_nr((rcur + lo_of(rcur, rp)) * 0.5);  // C on the x face below the point
#pragma unroll
  for (int m = 0; m < R2; ++m) {
    const int x = ta + 16 * m, xn = ta + 16 * ((m + 1) % R2);
    const double pn = P[rc + xn], rn = rho[rc + xn];
    const double pym = P[ryl + x], pyp = P[ryh + x], pzm = P[rzl + x], pzp = P[rzh + x];
    const double qym = rho[ryl + x], qyp = rho[ryh + x], qzm = rho[rzl + x], qzp = rho[rzh + x];
    const double r0 = rhs[rc + x];
    const double cx0n = rcp_nr((rn + lo_of(rn, rcur)) * 0.5);  // of the next point
    const double cx1 = hi_of(cx0, cx0n);                       // = C on the face below x + 1
    const double cy0 = rcp_nr((rcur + qym) * 0.5), cy1 = rcp_nr((qyp + rcur) * 0.5);
    const double cz0 = rcp_nr((rcur + qzm) * 0.5), cz1 = rcp_nr((qzp + rcur) * 0.5);
    v[m] = make_double2(residual_of(a.g, pcur, lo_of(pcur, pp), hi_of(pcur, pn), pym, pyp, pzm, pzp, cx0, cx1, cy0,
                                    cy1, cz0, cz1, r0, mean),
                        0.0);
    pp = pcur;
    pcur = pn;
    rcur = rn;
    cx0 = cx0n;
  }
}

// X forward pass: lpb consecutive x lines per block
struct k_xfwd {
  static constexpr int kBounds = kThreads;
  static __device__ __forceinline__ void run(const XArgs &a, int lpb) {
  if (*a.done != 0.0) return;  // the solve has converged
  extern __shared__ double2 lds[];
  const Geo &g = a.g;
  int nx = g.nx, mxh = nx / 2 + 1;
  long nlines = (long)g.ny * g.nz;
  Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
  double mean = a.scal[SC_SUM_RHS] / a.ncell;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  long line0 = (long)blk * lpb;
  for (int e = threadIdx.x; e < lpb * nx; e += block_dim_x()) {
    int l = e / nx, i = e - l * nx;
    long line = line0 + l;
    double r = 0.0;
    if (line < nlines) r = residual(a, ix, i, (int)(line % g.ny), (int)(line / g.ny), mean);
    lds[e] = make_double2(r, 0.0);
  }
  __syncthreads();
  lds_fft<false, false>(lds, nx, lpb, a.tw);
  for (int e = threadIdx.x; e < lpb * mxh; e += block_dim_x()) {
    int l = e / mxh, m = e - l * mxh;
    long line = line0 + l;
    if (line < nlines) a.spec[line * mxh + m] = lds[l * nx + m];
  }
}
};

struct XInvArgs {
  Geo g;
  double *P;
  const double2 *spec;
  const double2 *tw;
  double *partials;
  double ncell;
  const double *done;
  int overwrite = 0;  // three-pass form: P = dp (P unread, its own kernel instance) instead of P += dp
  double *Pout = nullptr;  // non-null: P + dp is written here and P is left as it was
  __device__ void to_global() {
    P = gptr(P);
    spec = gptr(spec);
    tw = gptr(tw);
    partials = gptr(partials);
    done = gptr(done);
    Pout = gptr(Pout);
  }
};

// X inverse pass: Hermitian extension, inverse FFT, P += dp, block max|dp|
struct k_xinv {
  static constexpr int kBounds = kThreads;
  static __device__ __forceinline__ void run(const XInvArgs &a, int lpb) {
  if (*a.done != 0.0) return;  // the solve has converged
  extern __shared__ double2 lds[];
  __shared__ double red[kThreads];
  const Geo &g = a.g;
  int nx = g.nx, mxh = nx / 2 + 1;
  long nlines = (long)g.ny * g.nz;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  long line0 = (long)blk * lpb;
  for (int e = threadIdx.x; e < lpb * mxh; e += block_dim_x()) {
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
  for (int e = threadIdx.x; e < lpb * nx; e += block_dim_x()) {
    int l = e / nx, i = e - l * nx;
    long line = line0 + l;
    if (line < nlines) {
      double dp = (-1) * lds[e].x / (a.ncell);
      const long o = line * nx + i;
      (a.Pout ? a.Pout : a.P)[o] = a.P[o] + dp;
      mx = fmax(mx, fabs(dp));
    }
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = block_dim_x() / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) a.partials[blk] = red[0];
}
};

// Y pass (forward or inverse) on a tile of kTile kx columns at fixed z
template <bool INV>
struct k_ypass {
  static constexpr int kBounds = kThreads;
  static __device__ __forceinline__ void run(double2 *spec, int mxh, int ny, int nz,
                                                    const double2 *__restrict__ tw, int kTile, const double *__restrict__ done) {
  if (*done != 0.0) return;  // the solve has converged
  extern __shared__ double2 lds[];
  int ntile = (mxh + kTile - 1) / kTile;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  int tile = blk % ntile, k = blk / ntile;
  int m0 = tile * kTile;
  int cols = min(kTile, mxh - m0);
  double2 *base = spec + (long)k * ny * mxh + m0;
  for (int e = threadIdx.x; e < ny * kTile; e += block_dim_x()) {
    int j = e / kTile, col = e - j * kTile;
    double2 v = col < cols ? base[(long)j * mxh + col] : make_double2(0.0, 0.0);
    lds[e] = v;  // element j of column col at j*kTile + col
  }
  __syncthreads();
  lds_fft<INV, true>(lds, ny, kTile, tw);
  for (int e = threadIdx.x; e < ny * kTile; e += block_dim_x()) {
    int j = e / kTile, col = e - j * kTile;
    if (col < cols) base[(long)j * mxh + col] = lds[e];
  }
}
};

// Z pass: forward FFT along z, divide by lambda (poisson.cpp:433-451), inverse
struct k_zpass {
  static constexpr int kBounds = kThreads;
  static __device__ __forceinline__ void run(double2 *spec, int mxh, int ny, int nz,
                                                    const double2 *__restrict__ tw,
                                                    const double *__restrict__ lamx,
                                                    const double *__restrict__ lamy,
                                                    const double *__restrict__ lamz, int kTile, const double *__restrict__ done) {
  if (*done != 0.0) return;  // the solve has converged
  extern __shared__ double2 lds[];
  int ntile = (mxh + kTile - 1) / kTile;
  unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  int tile = blk % ntile, j = blk / ntile;
  int m0 = tile * kTile;
  int cols = min(kTile, mxh - m0);
  long stride = (long)ny * mxh;
  double2 *base = spec + (long)j * mxh + m0;
  for (int e = threadIdx.x; e < nz * kTile; e += block_dim_x()) {
    int k = e / kTile, col = e - k * kTile;
    double2 v = col < cols ? base[(long)k * stride + col] : make_double2(0.0, 0.0);
    lds[e] = v;  // element k of column col at k*kTile + col
  }
  __syncthreads();
  lds_fft<false, true>(lds, nz, kTile, tw);
  for (int e = threadIdx.x; e < nz * kTile; e += block_dim_x()) {
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
  for (int e = threadIdx.x; e < nz * kTile; e += block_dim_x()) {
    int k = e / kTile, col = e - k * kTile;
    if (col < cols) base[(long)k * stride + col] = lds[e];
  }
}
};

// ---------------------------------------------------------------------------
// Four-step register FFT for lines of n = R1 * R2 points (R1 >= R2, both in
// {4, 8, 16}).  With i = t + R1 m and k = k2 + R2 k1,
//   X[k2 + R2 k1] = sum_t W_R1^{t k1} W_n^{t k2} sum_m x[t + R1 m] W_R2^{m k2}.
// Step A: the work item for t holds x[t + R1 m] (m < R2) in registers,
//         transforms over m and applies W_n^{t k2}.
// Exchange through LDS (one write, one barrier, one read).
// Step B: the work item for k2 gathers the R1 values over t and transforms
//         them: it ends holding X[k2 + R2 k1] for k1 < R1.
// Step B's output layout is step A's input layout of the transposed
// factorisation (R2, R1), so a forward transform can feed an inverse one
// without another exchange.
// ---------------------------------------------------------------------------
// e^{-2 pi i k/16}, k < 8, from cos/sin of multiples of pi/8
__device__ __forceinline__ double2 w16(int k) {
  constexpr double c1 = 0.9238795325112867, s1 = 0.3826834323650898, r = 0.7071067811865476;
  switch (k) {
    case 0: return make_double2(1.0, 0.0);
    case 1: return make_double2(c1, -s1);
    case 2: return make_double2(r, -r);
    case 3: return make_double2(s1, -c1);
    case 4: return make_double2(0.0, -1.0);
    case 5: return make_double2(-s1, -c1);
    case 6: return make_double2(-r, -r);
    default: return make_double2(-c1, -s1);
  }
}

// multiply by W_R^k (forward) or its conjugate (INV); k < R/2 is a constant
template <int R, bool INV>
__device__ __forceinline__ double2 twr(double2 a, int k) {
  if (k == 0) return a;
  if (4 * k == R) return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
  double2 w = w16(k * (16 / R));
  if (INV) w.y = -w.y;
  return cmul(a, w);
}

// in-register DFT of R points, natural order in and out (radix-2 DIT)
template <int R, bool INV>
struct RegDft {
  static __device__ __forceinline__ void run(double2 *v) {
    double2 e[R / 2], o[R / 2];
#pragma unroll
    for (int q = 0; q < R / 2; ++q) { e[q] = v[2 * q]; o[q] = v[2 * q + 1]; }
    RegDft<R / 2, INV>::run(e);
    RegDft<R / 2, INV>::run(o);
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
      double2 t = twr<R, INV>(o[q], q);
      v[q] = cadd(e[q], t);
      v[q + R / 2] = csub(e[q], t);
    }
  }
};
template <bool INV>
struct RegDft<1, INV> {
  static __device__ __forceinline__ void run(double2 *) {}
};

// W_16^m for any m (forward; INV conjugates)
template <bool INV>
__device__ __forceinline__ double2 w16m(int m) {
  m &= 15;
  double2 w = m < 8 ? w16(m) : make_double2(-w16(m - 8).x, -w16(m - 8).y);
  if (INV) w.y = -w.y;
  return w;
}

// a times -i (forward) or +i (INV)
template <bool INV>
__device__ __forceinline__ double2 rot_i(double2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// radix-4 butterfly in place on a[0..3] (natural order in and out)
template <bool INV>
__device__ __forceinline__ void dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3) {
  const double2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = rot_i<INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <bool INV>
struct RegDft<2, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool INV>
struct RegDft<4, INV> {
  static __device__ __forceinline__ void run(double2 *v) { dft4<INV>(v[0], v[1], v[2], v[3]); }
};

// 8 = 2 x 4: a radix-4 transform over n2 for each n1 (x[n1 + 2 n2]), the
// twiddles W_8^{n1 k1}, then radix-2 over n1: X[k1 + 4 k2]
template <bool INV>
struct RegDft<8, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    double2 y[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) y[q] = v[q];
    dft4<INV>(y[0], y[2], y[4], y[6]);  // n1 = 0: k1 at y[2 k1]
    dft4<INV>(y[1], y[3], y[5], y[7]);  // n1 = 1: k1 at y[1 + 2 k1]
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) y[1 + 2 * k1] = cmul(y[1 + 2 * k1], w16m<INV>(2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      v[k1] = cadd(y[2 * k1], y[1 + 2 * k1]);
      v[k1 + 4] = csub(y[2 * k1], y[1 + 2 * k1]);
    }
  }
};

// 16 = 4 x 4: radix-4 over n2 for each n1 (x[n1 + 4 n2]) in place, the
// twiddles W_16^{n1 k1}, radix-4 over n1 for each k1: X[k1 + 4 k2] is then
// in slot 4 k1 + k2, renamed into natural order
template <bool INV>
struct RegDft<16, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    double2 y[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) y[q] = v[q];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) dft4<INV>(y[n1], y[n1 + 4], y[n1 + 8], y[n1 + 12]);  // k1 at y[n1 + 4 k1]
#pragma unroll
    for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) y[n1 + 4 * k1] = cmul(y[n1 + 4 * k1], w16m<INV>(n1 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4<INV>(y[4 * k1], y[4 * k1 + 1], y[4 * k1 + 2], y[4 * k1 + 3]);  // k2 at + k2
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = y[4 * k1 + k2];
  }
};

// step A on values v[m] = x[t + R1 m]: transform over m, then W_n^{t k2}
template <int R2, bool INV>
__device__ __forceinline__ void step_a(double2 (&v)[R2], int t, const double2 *__restrict__ tw) {
  RegDft<R2, INV>::run(v);
#pragma unroll
  for (int k2 = 1; k2 < R2; ++k2) {
    double2 w = tw[t * k2];
    if (INV) w.y = -w.y;
    v[k2] = cmul(v[k2], w);
  }
}

// X forward, four-step: C lines per block, R1 work items per line in step A
template <int R1, int R2>
struct k_xfwd4 {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const XArgs &a, int C) {
  if (*a.done != 0.0) return;  // the solve has converged
  // The exchange moves real and imaginary parts in two rounds through an
  // LDS buffer of doubles: half the LDS per block, more blocks per CU.
  extern __shared__ double ldsd[];
  constexpr int n = R1 * R2, P = R1 + 1;  // padded row of the exchange
  const Geo &g = a.g;
  const int mxh = n / 2 + 1;
  const long nlines = (long)g.ny * g.nz;
  Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
  const double mean = a.scal[SC_SUM_RHS] / a.ncell;
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const long line0 = (long)blk * C;
  const int ta = threadIdx.x % R1, ca = threadIdx.x / R1;
  double2 v[R2];
  {
    const long line = line0 + ca;
    const bool live = line < nlines;
    const int j = (int)(line % g.ny), k = (int)(line / g.ny);
#pragma unroll
    for (int m = 0; m < R2; ++m) v[m] = make_double2(live ? residual(a, ix, ta + R1 * m, j, k, mean) : 0.0, 0.0);
    step_a<R2, false>(v, ta, a.tw);
  }
  const bool b_item = (int)threadIdx.x < C * R2;
  const int k2 = threadIdx.x % R2, cb = threadIdx.x / R2;
  double2 u[R1];
#pragma unroll
  for (int q = 0; q < R2; ++q) ldsd[(ca * R2 + q) * P + ta] = v[q].x;
  __syncthreads();
  if (b_item) {
#pragma unroll
    for (int t = 0; t < R1; ++t) u[t].x = ldsd[(cb * R2 + k2) * P + t];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R2; ++q) ldsd[(ca * R2 + q) * P + ta] = v[q].y;
  __syncthreads();
  if (b_item) {
    const long line = line0 + cb;
#pragma unroll
    for (int t = 0; t < R1; ++t) u[t].y = ldsd[(cb * R2 + k2) * P + t];
    RegDft<R1, false>::run(u);
    if (line < nlines) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        int m = k2 + R2 * k1;
        if (m < mxh) a.spec[line * mxh + m] = u[k1];
      }
    }
  }
}
};

// X inverse, four-step: Hermitian extension, P += dp, block max |dp|
template <int R1, int R2>
struct k_xinv4 {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const XInvArgs &a, int C) {
  if (*a.done != 0.0) return;  // the solve has converged
  extern __shared__ double2 lds[];
  __shared__ double red[256];
  constexpr int n = R1 * R2, P = R1 + 1;
  const Geo &g = a.g;
  const int mxh = n / 2 + 1;
  const long nlines = (long)g.ny * g.nz;
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const long line0 = (long)blk * C;
  {
    const int t = threadIdx.x % R1, c = threadIdx.x / R1;
    const long line = line0 + c;
    double2 v[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      int i = t + R1 * m;
      double2 s = make_double2(0.0, 0.0);
      if (line < nlines) {
        if (i < mxh) {
          s = a.spec[line * mxh + i];
          if (i == 0 || i == n / 2) s.y = 0.0;
        } else {
          s = a.spec[line * mxh + (n - i)];
          s.y = -s.y;
        }
      }
      v[m] = s;
    }
    step_a<R2, true>(v, t, a.tw);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) lds[(c * R2 + k2) * P + t] = v[k2];
  }
  __syncthreads();
  double mx = 0.0;
  if ((int)threadIdx.x < C * R2) {
    const int k2 = threadIdx.x % R2, c = threadIdx.x / R2;
    const long line = line0 + c;
    double2 u[R1];
#pragma unroll
    for (int t = 0; t < R1; ++t) u[t] = lds[(c * R2 + k2) * P + t];
    RegDft<R1, true>::run(u);
    if (line < nlines) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        double dp = (-1) * u[k1].x / (a.ncell);
        const long o = line * n + k2 + R2 * k1;
        (a.Pout ? a.Pout : a.P)[o] = a.P[o] + dp;
        mx = fmax(mx, fabs(dp));
      }
    }
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = block_dim_x() / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) a.partials[blk] = red[0];
}
};

// Strided lines (the y and z passes): C kx columns per block, consecutive
// threads on consecutive columns.  `base` points at element 0 of column 0,
// `stride` is the distance between consecutive line elements.
template <int R1, int R2, bool INV>
__device__ __forceinline__ void strided_fwd_half(double2 (&v)[R2], double2 *lds, const double2 *base, long stride,
                                                 int C, int cols, const double2 *__restrict__ tw) {
  const int col = threadIdx.x % C, t = threadIdx.x / C;
#pragma unroll
  for (int m = 0; m < R2; ++m)
    v[m] = col < cols ? base[(long)(t + R1 * m) * stride + col] : make_double2(0.0, 0.0);
  step_a<R2, INV>(v, t, tw);
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) lds[(k2 * R1 + t) * C + col] = v[k2];
}

template <int R1, int R2, bool INV>
struct k_ypass4 {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(double2 *spec, int mxh, int ny, int nz,
                                                const double2 *__restrict__ tw, int C,
                                                const double *__restrict__ done) {
  if (*done != 0.0) return;  // the solve has converged
  extern __shared__ double2 lds[];
  const int ntile = (mxh + C - 1) / C;
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int tile = blk % ntile, k = blk / ntile;
  const int m0 = tile * C, cols = min(C, mxh - m0);
  double2 *base = spec + (long)k * ny * mxh + m0;
  {
    double2 v[R2];
    strided_fwd_half<R1, R2, INV>(v, lds, base, mxh, C, cols, tw);
  }
  __syncthreads();
  if ((int)threadIdx.x < C * R2) {
    const int col = threadIdx.x % C, k2 = threadIdx.x / C;
    double2 u[R1];
#pragma unroll
    for (int t = 0; t < R1; ++t) u[t] = lds[(k2 * R1 + t) * C + col];
    RegDft<R1, INV>::run(u);
    if (col < cols) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) base[(long)(k2 + R2 * k1) * mxh + col] = u[k1];
    }
  }
}
};

// Z pass: forward FFT along z, divide by lambda (poisson.cpp:433-451), inverse
// FFT along z with the transposed factorisation, all on one load and store.
template <int R1, int R2>
struct k_zpass4 {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(double2 *spec, int mxh, int ny, int nz,
                                                const double2 *__restrict__ tw,
                                                const double *__restrict__ lamx,
                                                const double *__restrict__ lamy,
                                                const double *__restrict__ lamz, int C,
                                                const double *__restrict__ done) {
  if (*done != 0.0) return;  // the solve has converged
  extern __shared__ double2 lds[];
  const int ntile = (mxh + C - 1) / C;
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int tile = blk % ntile, j = blk / ntile;
  const int m0 = tile * C, cols = min(C, mxh - m0);
  const long stride = (long)ny * mxh;
  double2 *base = spec + (long)j * mxh + m0;
  {
    double2 v[R2];
    strided_fwd_half<R1, R2, false>(v, lds, base, stride, C, cols, tw);
  }
  __syncthreads();
  const bool b_item = (int)threadIdx.x < C * R2;
  const int col = threadIdx.x % C;
  double2 u[R1];
  if (b_item) {
    const int k2 = threadIdx.x / C;
#pragma unroll
    for (int t = 0; t < R1; ++t) u[t] = lds[(k2 * R1 + t) * C + col];
    RegDft<R1, false>::run(u);
    // u[k1] = X[k2 + R2 k1]: divide by lambda except the global zero mode
    const int m = m0 + col;
    const double lxy = lamx[m < mxh ? m : 0] + lamy[j];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      int kz = k2 + R2 * k1;
      if (m + j + kz > 0) {
        double lam = lxy + lamz[kz];
        u[k1].y /= lam;
        u[k1].x /= lam;
      }
    }
    // inverse step A with factorisation (R2, R1): the work item for t' = k2
    // holds x[t' + R2 m'] = u[m']
    RegDft<R1, true>::run(u);
#pragma unroll
    for (int q = 1; q < R1; ++q) {
      double2 w = tw[k2 * q];
      w.y = -w.y;
      u[q] = cmul(u[q], w);
    }
  }
  __syncthreads();  // the forward exchange has been read
  if (b_item) {
    const int k2 = threadIdx.x / C;
#pragma unroll
    for (int q = 0; q < R1; ++q) lds[(q * R2 + k2) * C + col] = u[q];
  }
  __syncthreads();
  {
    // inverse step B: the work item for q < R1 gathers over t' < R2
    const int q = threadIdx.x / C;
    double2 w2[R2];
#pragma unroll
    for (int t = 0; t < R2; ++t) w2[t] = lds[(q * R2 + t) * C + col];
    RegDft<R2, true>::run(w2);
    if (col < cols) {
#pragma unroll
      for (int k1 = 0; k1 < R2; ++k1) base[(long)(q + R1 * k1) * stride + col] = w2[k1];
    }
  }
}
};

// ---------------------------------------------------------------------------
// Three-pass iteration (every axis length 16..256).  The y transform is split
// between the passes along its four-step factorisation ny = Y1 Y2, with
// y = t + Y1 m:
//   pass 1, block (z, t):   residual and x transform of the Y2 rows
//                           y = t + Y1 m, then y step A for every kx
//                           (DFT over m, twiddle W_ny^{t k2});
//   pass 2, block (kx, k2): y step B (DFT over t, ky = k2 + Y2 k1), z
//                           transform, divide by lambda, inverse z, inverse
//                           y step B (over k1 back to t);
//   pass 3, block (z, t):   inverse y step A (twiddle W_ny^{-t k2}, DFT over
//                           k2 back to m) for every kx, inverse x transform
//                           of the Y2 rows, P += dp, block max |dp|.
// Between the passes the spectrum is stored as
//   S[(k2 mxh + kx) CS + z Y1 + t],   CS = nz Y1 + 8,
// so a pass-2 block reads and writes one contiguous nz Y1 run, and the Y1
// blocks (z, 0..Y1-1) of passes 1 and 3, which run next to each other on one
// XCD, share every 256-byte run of it.  In passes 1 and 3 the lanes of a wave
// take consecutive kx, 16 nz Y1 + 128 bytes apart: the 128 spare bytes per
// run spread those accesses over the L2 channels (at a stride of a power of
// two they would all queue on one).  Two passes over the spectrum fewer than
// the five-pass form.
constexpr int kRunPad = 8;  // spare complex values after each nz Y1 run
// ---------------------------------------------------------------------------
template <int R1, int R2, int Y2>
struct k_xy_fwd {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const XArgs &a, int Y1, const double2 *__restrict__ twy) {
  if (*a.done != 0.0) return;  // the solve has converged
  extern __shared__ double ldsd[];
  constexpr int n = R1 * R2, P = R1 + 1, SP = Y2 + 1;  // padded rows of the x exchange and of the kx columns
  const Geo &g = a.g;
  const int mxh = n / 2 + 1;
  Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
  const double mean = a.scal[SC_SUM_RHS] / a.ncell;
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int t = (int)(blk % (unsigned)Y1), z = (int)(blk / (unsigned)Y1);
  // x step A: R1 work items per row, row ca is y = t + Y1 ca
  const int ta = threadIdx.x % R1, ca = threadIdx.x / R1;
  double2 v[R2];
  {
    const int j = t + Y1 * ca;
    if (a.in) {
      const double *row = a.in + (long)z * g.nxy + (long)j * g.nx + ta;
#pragma unroll
      for (int m = 0; m < R2; ++m) v[m] = make_double2(row[R1 * m], 0.0);
    } else if constexpr (R1 == 16) {
      row_residuals<R2>(a, ix, ta, j, z, mean, v);
    } else {
#pragma unroll
      for (int m = 0; m < R2; ++m) v[m] = make_double2(residual(a, ix, ta + R1 * m, j, z, mean), 0.0);
    }
    step_a<R2, false>(v, ta, a.tw);
  }
  const bool b_item = (int)threadIdx.x < Y2 * R2;
  const int k2x = threadIdx.x % R2, cb = threadIdx.x / R2;
  double2 u[R1];
#pragma unroll
  for (int q = 0; q < R2; ++q) ldsd[(ca * R2 + q) * P + ta] = v[q].x;
  __syncthreads();
  if (b_item) {
#pragma unroll
    for (int q = 0; q < R1; ++q) u[q].x = ldsd[(cb * R2 + k2x) * P + q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R2; ++q) ldsd[(ca * R2 + q) * P + ta] = v[q].y;
  __syncthreads();
  if (b_item) {
#pragma unroll
    for (int q = 0; q < R1; ++q) u[q].y = ldsd[(cb * R2 + k2x) * P + q];
    RegDft<R1, false>::run(u);
  }
  __syncthreads();  // the exchange buffer now holds the rows' half spectra
  double2 *S = reinterpret_cast<double2 *>(ldsd);  // S[kx SP + row]
  if (b_item) {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int kx = k2x + R2 * k1;
      if (kx < mxh) S[kx * SP + cb] = u[k1];
    }
  }
  __syncthreads();
  // y step A for every kx
  const long cs = (long)g.nz * Y1 + kRunPad;
  const long ks = cs * mxh;  // distance between consecutive k2
  for (int kx = threadIdx.x; kx < mxh; kx += block_dim_x()) {
    double2 w[Y2];
#pragma unroll
    for (int m = 0; m < Y2; ++m) w[m] = S[kx * SP + m];
    step_a<Y2, false>(w, t, twy);
    double2 *dst = a.spec + kx * cs + z * Y1 + t;
#pragma unroll
    for (int k2 = 0; k2 < Y2; ++k2) dst[k2 * ks] = w[k2];
  }
}
};

template <int Y1, int Z1, int Z2>
struct k_yz_pass {
  // one lane an item of the largest stage, whole waves: 64 lanes at nz = 64
  // and Y1 = 8 (the 64^3 boxes), where 256 would leave three waves of four idle
  // in every transform stage
  static constexpr int kItems = std::max(Z1 * Z2, Y1 * std::max(Z1, Z2));
  static constexpr int kBounds = kItems <= 64 ? 64 : (kItems <= 128 ? 128 : 256);
  static __device__ __forceinline__ void run(double2 *spec, int mxh, int Y2, const double2 *__restrict__ twz,
                                                 const double *__restrict__ lamx,
                                                 const double *__restrict__ lamy,
                                                 const double *__restrict__ lamz,
                                                 const double *__restrict__ done) {
  if (*done != 0.0) return;  // the solve has converged
  extern __shared__ double2 L[];
  constexpr int nz = Z1 * Z2, NE = nz * Y1, RS = nz + nz / 16;
  auto pad = [](int e) { return e + (e >> 4); };  // one spare slot every 16: fewer bank conflicts
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int k2 = (int)(blk / (unsigned)mxh), kx = (int)(blk % (unsigned)mxh);
  // element (z, t) at z Y1 + t.  The run and the tables are read through
  // global-address-space pointers: in the batched launch the arguments come
  // from a table in memory as generic pointers, whose flat loads also count
  // against the LDS wait counter (and the register staging below went
  // through scratch); a kernel's own pointer arguments are known to be global.
  typedef double dv2 __attribute__((ext_vector_type(2)));  // a builtin pair: copied through any address space
  typedef __attribute__((address_space(1))) dv2 gdv2;
  typedef __attribute__((address_space(1))) const dv2 gcdv2;
  typedef __attribute__((address_space(1))) const double gcdouble;
  gdv2 *base = (gdv2 *)(spec + (long)blk * (NE + kRunPad));
  const int tid = threadIdx.x;
  // the z twiddles and lambda_z beside the run (yz_shared_bytes), copied in
  // with it: the transform stages then wait on LDS for them, not on global
  // memory (74 against 98 us a pass at 256^3); lambda_x + lambda_y of the item
  // this thread takes in z step B, read up front
  double2 *twl = L + Y1 * RS;
  double *lzl = reinterpret_cast<double *>(twl + nz);
  const double lxy = tid < Y1 * Z2 ? ((gcdouble *)lamx)[kx] + ((gcdouble *)lamy)[k2 + Y2 * (tid / Z2)] : 0.0;
  for (int e = tid; e < nz; e += kBounds) {
    const dv2 w = ((gcdv2 *)twz)[e];
    twl[e] = make_double2(w.x, w.y);
    lzl[e] = ((gcdouble *)lamz)[e];
  }
  // the run into LDS: every load of a thread issued before the first LDS
  // write (a loop over the block width read at run time waited on each load
  // in turn); yz_launch starts kBounds threads a block
  constexpr int NU = (NE + kBounds - 1) / kBounds;
  {
    double2 r[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int e = tid + u * kBounds;
      if (NE % kBounds == 0 || e < NE) {
        const dv2 q = base[e];
        r[u] = make_double2(q.x, q.y);
      }
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int e = tid + u * kBounds;
      if (NE % kBounds == 0 || e < NE) L[pad(e)] = r[u];
    }
  }
  __syncthreads();
  // y step B: item z, DFT over t
  const bool zi = tid < nz;
  double2 u[Y1];
  if (zi) {
#pragma unroll
    for (int q = 0; q < Y1; ++q) u[q] = L[pad(tid * Y1 + q)];
    RegDft<Y1, false>::run(u);  // u[k1] at ky = k2 + Y2 k1
  }
  __syncthreads();
  if (zi) {
#pragma unroll
    for (int k1 = 0; k1 < Y1; ++k1) L[k1 * RS + pad(tid)] = u[k1];
  }
  __syncthreads();
  // z step A: item (k1, tz), over z = tz + Z1 mz
  const bool ai = tid < Y1 * Z1;
  const int tz = tid % Z1, ka = tid / Z1;
  {
    double2 v[Z2];
    if (ai) {
#pragma unroll
      for (int m = 0; m < Z2; ++m) v[m] = L[ka * RS + pad(tz + Z1 * m)];
      step_a<Z2, false>(v, tz, twl);
    }
    __syncthreads();
    if (ai) {
#pragma unroll
      for (int q = 0; q < Z2; ++q) L[ka * RS + pad(q * Z1 + tz)] = v[q];
    }
    __syncthreads();
  }
  // z step B: item (k1, kz2), DFT over tz; divide by lambda; inverse step A
  // with the transposed factorisation
  const bool bi = tid < Y1 * Z2;
  const int kz2 = tid % Z2, kb = tid / Z2;
  {
    double2 w[Z1];
    if (bi) {
#pragma unroll
      for (int q = 0; q < Z1; ++q) w[q] = L[kb * RS + pad(kz2 * Z1 + q)];
      RegDft<Z1, false>::run(w);  // w[kz1] at kz = kz2 + Z2 kz1
      const int ky = k2 + Y2 * kb;
#pragma unroll
      for (int kz1 = 0; kz1 < Z1; ++kz1) {
        const int kz = kz2 + Z2 * kz1;
        if (kx + ky + kz > 0) {
          const double lam = lxy + lzl[kz];
          w[kz1].y /= lam;
          w[kz1].x /= lam;
        }
      }
      RegDft<Z1, true>::run(w);
#pragma unroll
      for (int q = 1; q < Z1; ++q) {
        double2 tw = twl[kz2 * q];
        tw.y = -tw.y;
        w[q] = cmul(w[q], tw);
      }
    }
    __syncthreads();
    if (bi) {
#pragma unroll
      for (int q = 0; q < Z1; ++q) L[kb * RS + pad(q * Z2 + kz2)] = w[q];
    }
    __syncthreads();
  }
  // inverse z step B: item (k1, q), DFT over kz2, z = q + Z1 k
  {
    const int q = tid % Z1, kc = tid / Z1;
    double2 x[Z2];
    if (ai) {
#pragma unroll
      for (int s = 0; s < Z2; ++s) x[s] = L[kc * RS + pad(q * Z2 + s)];
      RegDft<Z2, true>::run(x);
    }
    __syncthreads();
    if (ai) {
#pragma unroll
      for (int s = 0; s < Z2; ++s) L[kc * RS + pad(q + Z1 * s)] = x[s];
    }
    __syncthreads();
  }
  // inverse y step B: item z, DFT over k1 back to t
  if (zi) {
#pragma unroll
    for (int k1 = 0; k1 < Y1; ++k1) u[k1] = L[k1 * RS + pad(tid)];
    RegDft<Y1, true>::run(u);
  }
  __syncthreads();
  if (zi) {
#pragma unroll
    for (int q = 0; q < Y1; ++q) L[pad(tid * Y1 + q)] = u[q];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int e = tid + u * kBounds;
    if (NE % kBounds == 0 || e < NE) {
      const double2 q = L[pad(e)];
      base[e] = dv2{q.x, q.y};
    }
  }
}
};

// ---------------------------------------------------------------------------
// Pass 1 with the x transform of the real residual done as a complex
// transform of half the length (x lengths 2 M = 64, 128, 256; 16 lanes a row):
// lane ta of a row holds the pairs z[j] = x[2j] + i x[2j+1], j = ta + 16 m
// (m < R2H, M = 16 R2H), loaded with 16-byte accesses.  The M-point four-step
// transform Z of z gives the half spectrum of x by
//   X[k] = E[k] + W_2M^k O[k],  E = (Z[k] + conj Z[M-k]) / 2,
//                               O = (Z[k] - conj Z[M-k]) / 2i   (k <= M).
// Half the butterflies and half the load instructions of the full-length
// complex transform of a real row; the residual's x neighbours come from the
// adjacent lanes of the row by DPP, as in the full-length form.
// ---------------------------------------------------------------------------

// the residuals of one x-row in the pair layout: v[m] = (r(x0), r(x0 + 1)),
// x0 = 2 ta + 32 m.  The x neighbours outside the pair come from the lanes
// beside it by DPP row rotation, as in row_residuals; the same arithmetic.
template <int R2H>
__device__ __forceinline__ void row_residuals_pairs(const XArgs &a, const Ix &ix, int ta, int j, int z, double mean,
                                                    double2 (&v)[R2H]) {
  const double *__restrict__ rho = a.rho;
  const double *__restrict__ P = a.P;
  const double *__restrict__ rhs = a.rhs;
  const long rc = (long)(z & ix.mz) * ix.nxy + (long)(j & ix.my) * ix.nx;
  const long ryl = (long)(z & ix.mz) * ix.nxy + (long)((j - 1) & ix.my) * ix.nx;
  const long ryh = (long)(z & ix.mz) * ix.nxy + (long)((j + 1) & ix.my) * ix.nx;
  const long rzl = (long)((z - 1) & ix.mz) * ix.nxy + (long)(j & ix.my) * ix.nx;
  const long rzh = (long)((z + 1) & ix.mz) * ix.nxy + (long)(j & ix.my) * ix.nx;
  auto ld2 = [&](const double *f, long row, int m) {
    return *reinterpret_cast<const double2 *>(f + row + 2 * ta + 32 * m);
  };
  // x0 - 1: the odd point of the lane before (lane 15 of the pair before, at ta 0)
  auto lo_of = [&](double cur_odd, double prev_odd) {
    const double u = row_rotate<0x121>(cur_odd), w = row_rotate<0x121>(prev_odd);
    return ta == 0 ? w : u;
  };
  // x0 + 2: the even point of the lane after (lane 0 of the pair after, at ta 15)
  auto hi_of = [&](double cur_even, double next_even) {
    const double u = row_rotate<0x12F>(cur_even), w = row_rotate<0x12F>(next_even);
    return ta == 15 ? w : u;
  };
  double2 pp = ld2(P, rc, R2H - 1), rp = ld2(rho, rc, R2H - 1);  // the pair before m = 0
  double2 pc = ld2(P, rc, 0), rcc = ld2(rho, rc, 0);
  double cxa = rcp_nr((rcc.x + lo_of(rcc.y, rp.y)) * 0.5);  // C on the face below x0
#pragma unroll
  for (int m = 0; m < R2H; ++m) {
    const int mn = (m + 1) % R2H;
    const double2 pn = ld2(P, rc, mn), rn = ld2(rho, rc, mn);
    const double2 pym = ld2(P, ryl, m), pyp = ld2(P, ryh, m), pzm = ld2(P, rzl, m), pzp = ld2(P, rzh, m);
    const double2 qym = ld2(rho, ryl, m), qyp = ld2(rho, ryh, m), qzm = ld2(rho, rzl, m), qzp = ld2(rho, rzh, m);
    const double2 r0 = ld2(rhs, rc, m);
    const double cxb = rcp_nr((rcc.y + rcc.x) * 0.5);                    // the face between x0 and x0 + 1
    const double cxa_n = rcp_nr((rn.x + lo_of(rn.y, rcc.y)) * 0.5);      // below x0 of the next pair
    const double cxc = hi_of(cxa, cxa_n);                                // above x0 + 1
    const double cy0a = rcp_nr((rcc.x + qym.x) * 0.5), cy1a = rcp_nr((qyp.x + rcc.x) * 0.5);
    const double cz0a = rcp_nr((rcc.x + qzm.x) * 0.5), cz1a = rcp_nr((qzp.x + rcc.x) * 0.5);
    const double cy0b = rcp_nr((rcc.y + qym.y) * 0.5), cy1b = rcp_nr((qyp.y + rcc.y) * 0.5);
    const double cz0b = rcp_nr((rcc.y + qzm.y) * 0.5), cz1b = rcp_nr((qzp.y + rcc.y) * 0.5);
    const double ra = residual_of(a.g, pc.x, lo_of(pc.y, pp.y), pc.y, pym.x, pyp.x, pzm.x, pzp.x, cxa, cxb, cy0a,
                                  cy1a, cz0a, cz1a, r0.x, mean);
    const double rb = residual_of(a.g, pc.y, pc.x, hi_of(pc.x, pn.x), pym.y, pyp.y, pzm.y, pzp.y, cxb, cxc, cy0b,
                                  cy1b, cz0b, cz1b, r0.y, mean);
    v[m] = make_double2(ra, rb);
    pp = pc;
    pc = pn;
    rcc = rn;
    cxa = cxa_n;
  }
}

// pass 1, packed: residual rows, the M-point x transform, the half spectrum
// recovered on the way out of LDS, y step A
template <int R2H, int Y2>
struct k_xy_fwd_rp {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const XArgs &a, int Y1, const double2 *__restrict__ twy) {
    if (*a.done != 0.0) return;  // the solve has converged
    extern __shared__ double ldsd[];
    constexpr int R1 = 16, M = R1 * R2H, P = R1 + 1, SP = Y2 + 1;
    const Geo &g = a.g;
    const int mxh = M + 1;
    Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
    const double mean = a.scal[SC_SUM_RHS] / a.ncell;
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int t = (int)(blk % (unsigned)Y1), z = (int)(blk / (unsigned)Y1);
    const int ta = threadIdx.x % R1, ca = threadIdx.x / R1;
    double2 v[R2H];
    if (a.in) {
      const double *row = a.in + (long)z * g.nxy + (long)(t + Y1 * ca) * g.nx + 2 * ta;
#pragma unroll
      for (int m = 0; m < R2H; ++m) v[m] = *reinterpret_cast<const double2 *>(row + 32 * m);
    } else {
      row_residuals_pairs<R2H>(a, ix, ta, t + Y1 * ca, z, mean, v);
    }
    // step A of the M-point transform: over m, then W_M^{ta k2} = W_2M^{2 ta k2}
    RegDft<R2H, false>::run(v);
#pragma unroll
    for (int k2 = 1; k2 < R2H; ++k2) v[k2] = cmul(v[k2], a.tw[2 * ta * k2]);
    const bool b_item = (int)threadIdx.x < Y2 * R2H;
    const int k2x = threadIdx.x % R2H, cb = threadIdx.x / R2H;
    double2 u[R1];
#pragma unroll
    for (int q = 0; q < R2H; ++q) ldsd[(ca * R2H + q) * P + ta] = v[q].x;
    __syncthreads();
    if (b_item) {
#pragma unroll
      for (int q = 0; q < R1; ++q) u[q].x = ldsd[(cb * R2H + k2x) * P + q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R2H; ++q) ldsd[(ca * R2H + q) * P + ta] = v[q].y;
    __syncthreads();
    if (b_item) {
#pragma unroll
      for (int q = 0; q < R1; ++q) u[q].y = ldsd[(cb * R2H + k2x) * P + q];
      RegDft<R1, false>::run(u);  // u[k1] = Z[k2x + R2H k1] of row cb
    }
    __syncthreads();
    double2 *S = reinterpret_cast<double2 *>(ldsd);  // S[k SP + row], k < M
    if (b_item) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) S[(k2x + R2H * k1) * SP + cb] = u[k1];
    }
    __syncthreads();
    const long cs = (long)g.nz * Y1 + kRunPad;
    const long ks = cs * mxh;
    for (int kx = threadIdx.x; kx < mxh; kx += block_dim_x()) {
      const double2 wn = a.tw[kx];  // W_2M^kx
      const int ka = kx & (M - 1), kb = (M - kx) & (M - 1);
      double2 w[Y2];
#pragma unroll
      for (int m = 0; m < Y2; ++m) {
        const double2 A = S[ka * SP + m], B = S[kb * SP + m];
        const double2 E = make_double2((A.x + B.x) * 0.5, (A.y - B.y) * 0.5);   // (A + conj B) / 2
        const double2 D = make_double2((A.x - B.x) * 0.5, (A.y + B.y) * 0.5);   // (A - conj B) / 2
        const double2 O = make_double2(D.y, -D.x);                              // D / i
        w[m] = cadd(E, cmul(wn, O));
      }
      step_a<Y2, false>(w, t, twy);
      double2 *dst = a.spec + kx * cs + z * Y1 + t;
#pragma unroll
      for (int k2 = 0; k2 < Y2; ++k2) dst[k2 * ks] = w[k2];
    }
  }
};

template <int R1, int R2, int Y2, bool Overwrite = false, bool OutOfPlace = false>
struct k_xy_inv {
  static constexpr int kBounds = 256;
  static __device__ __forceinline__ void run(const XInvArgs &a, int Y1, const double2 *__restrict__ twy) {
  if (*a.done != 0.0) return;  // the solve has converged
  extern __shared__ double ldsd[];
  __shared__ double red[256];
  constexpr int n = R1 * R2, P = R1 + 1, SP = Y2 + 1;
  const Geo &g = a.g;
  const int mxh = n / 2 + 1;
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int t = (int)(blk % (unsigned)Y1), z = (int)(blk / (unsigned)Y1);
  double2 *S = reinterpret_cast<double2 *>(ldsd);  // S[kx SP + row]
  // inverse y step A for every kx: row m is y = t + Y1 m
  const long cs = (long)g.nz * Y1 + kRunPad;
  const long ks = cs * mxh;
  for (int kx = threadIdx.x; kx < mxh; kx += block_dim_x()) {
    const double2 *src = a.spec + kx * cs + z * Y1 + t;
    double2 w[Y2];
#pragma unroll
    for (int k2 = 0; k2 < Y2; ++k2) {
      w[k2] = src[k2 * ks];
      if (k2 > 0) {
        double2 tw = twy[t * k2];
        tw.y = -tw.y;
        w[k2] = cmul(w[k2], tw);
      }
    }
    RegDft<Y2, true>::run(w);
#pragma unroll
    for (int m = 0; m < Y2; ++m) S[kx * SP + m] = w[m];
  }
  __syncthreads();
  // inverse x: Hermitian extension of row c, step A over i = tx + R1 mm
  const int tx = threadIdx.x % R1, c = threadIdx.x / R1;
  double2 v[R2];
#pragma unroll
  for (int m = 0; m < R2; ++m) {
    const int i = tx + R1 * m;
    double2 s;
    if (i < mxh) {
      s = S[i * SP + c];
      if (i == 0 || i == n / 2) s.y = 0.0;
    } else {
      s = S[(n - i) * SP + c];
      s.y = -s.y;
    }
    v[m] = s;
  }
  step_a<R2, true>(v, tx, a.tw);
  __syncthreads();  // S is reused by the exchange
  const bool b_item = (int)threadIdx.x < Y2 * R2;
  const int k2x = threadIdx.x % R2, cb = threadIdx.x / R2;
  double2 u[R1];
#pragma unroll
  for (int q = 0; q < R2; ++q) ldsd[(c * R2 + q) * P + tx] = v[q].x;
  __syncthreads();
  if (b_item) {
#pragma unroll
    for (int q = 0; q < R1; ++q) u[q].x = ldsd[(cb * R2 + k2x) * P + q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R2; ++q) ldsd[(c * R2 + q) * P + tx] = v[q].y;
  __syncthreads();
  double mx = 0.0;
  if (b_item) {
#pragma unroll
    for (int q = 0; q < R1; ++q) u[q].y = ldsd[(cb * R2 + k2x) * P + q];
    RegDft<R1, true>::run(u);
    const long l0 = ((long)z * g.ny + t + Y1 * cb) * n;
    if constexpr (OutOfPlace) {
      const double *__restrict__ Pin = a.P + l0;
      double *__restrict__ Po = a.Pout + l0;
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        double dp = (-1) * u[k1].x / (a.ncell);
        Po[k2x + R2 * k1] = Pin[k2x + R2 * k1] + dp;
        mx = fmax(mx, fabs(dp));
      }
    } else {
      double *Pl = a.P + l0;
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        double dp = (-1) * u[k1].x / (a.ncell);
        double *Pp = Pl + k2x + R2 * k1;
        if constexpr (Overwrite) *Pp = dp;
        else *Pp = *Pp + dp;
        mx = fmax(mx, fabs(dp));
      }
    }
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = block_dim_x() / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) a.partials[blk] = red[0];
}
};

// End of one iteration: max |dp| over the block maxima into SC_MAX_DP, the
// iteration count, and the loop test of poisson.cpp:391-462 (continue while
// count < Iteration and max |dp| > epsilon) into SC_DONE.  A skipped
// iteration leaves every slot alone.
// The residual of poisson::FFT with the caller's face coefficient C
// (poisson.cpp:391-411): r = (div_F2C(C grad_C2F P) - RHS) cc with
// cc = -1/I_F2C(C), the mean of the cell's six face values; RHS is used as
// given.  One thread per cell; C.x at the low x face of its cell.
struct k_residual_coef {
  static constexpr int kBounds = kThreads;
  static __device__ __forceinline__ void run(Geo g, const double *__restrict__ cx, const double *__restrict__ cy,
                                             const double *__restrict__ cz, const double *__restrict__ P,
                                             const double *__restrict__ rhs, double *__restrict__ res,
                                             const double *done) {
    if (*done != 0.0) return;
    const long e = (long)blockIdx.x * block_dim_x() + threadIdx.x;
    if (e >= g.n) return;
    const int i = (int)(e % g.nx), j = (int)((e / g.nx) % g.ny), k = (int)(e / g.nxy);
    Ix ix{g.nx - 1, g.ny - 1, g.nz - 1, g.nx, g.nxy};
    res[e] = residual_of(g, P[e], P[ix(i - 1, j, k)], P[ix(i + 1, j, k)], P[ix(i, j - 1, k)], P[ix(i, j + 1, k)],
                         P[ix(i, j, k - 1)], P[ix(i, j, k + 1)], cx[e], cx[ix(i + 1, j, k)], cy[e],
                         cy[ix(i, j + 1, k)], cz[e], cz[ix(i, j, k + 1)], rhs[e], 0.0);
  }
};

struct k_finish_iter {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(const double *__restrict__ partials, long nparts,
                                                      double *scal, int iteration, double epsilon) {
  __shared__ double sh[1024];
  if (scal[SC_DONE] != 0.0) return;
  double v = 0.0;
  for (long q = threadIdx.x; q < nparts; q += 1024) v = fmax(v, partials[q]);  // coalesced
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double err = sh[0], it = scal[SC_ITERS] + 1.0;
    scal[SC_MAX_DP] = err;
    scal[SC_ITERS] = it;
    scal[SC_DONE] = (it < iteration && err > epsilon) ? 0.0 : 1.0;
  }
}
};

// Start of a solve: no iteration yet, err = 100 (poisson.cpp:391-396)
struct k_solve_begin {
  static constexpr int kBounds = 1024;
  static __device__ __forceinline__ void run(double *scal, int iteration, double epsilon) {
  scal[SC_ITERS] = 0.0;
  scal[SC_DONE] = (0 < iteration && 100.0 > epsilon) ? 0.0 : 1.0;
}
};

// the loop state of a solve (iterations run, stopped) into dst[0..1]
struct k_solve_state {
  static constexpr int kBounds = 64;
  static __device__ __forceinline__ void run(const double *scal, double *dst) {
    if (threadIdx.x == 0) {
      dst[0] = scal[SC_ITERS];
      dst[1] = scal[SC_DONE];
    }
  }
};

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
  // the three-pass layout leaves kRunPad spare values after each of its mxh Y2 runs
  if (hipMalloc(&c->spec[0], sizeof(double2) * mxh * ((long)g.ny * g.nz + 16 * kRunPad)) != hipSuccess)
    return TPX_E_OOM;
  return TPX_OK;
}

// (R1, R2) of the four-step kernels for an axis length, or false when the
// length uses the LDS Stockham kernels (8 and 512)
static bool four_step(int n, int &r1, int &r2) {
  switch (n) {
    case 16: r1 = 4; r2 = 4; return true;
    case 32: r1 = 8; r2 = 4; return true;
    case 64: r1 = 8; r2 = 8; return true;
    case 128: r1 = 16; r2 = 8; return true;
    case 256: r1 = 16; r2 = 16; return true;
    default: return false;
  }
}

#define TPX_FOUR_STEP_SWITCH(n, LAUNCH) \
  switch (n) {                           \
    case 16: LAUNCH(4, 4); break;        \
    case 32: LAUNCH(8, 4); break;        \
    case 64: LAUNCH(8, 8); break;        \
    case 128: LAUNCH(16, 8); break;      \
    default: LAUNCH(16, 16); break;      \
  }

// launch helpers of the three-pass iteration: template arguments from the
// run-time factorisations
template <int R1, int R2>
void xy_fwd_launch(unsigned nb, size_t sh, hipStream_t st, const XArgs &xa, int Y1, int Y2, const double2 *twy) {
  switch (Y2) {
    case 4: tpx_launch<k_xy_fwd<R1, R2, 4>>(nb, 4 * R1, sh, st, xa, Y1, twy); break;
    case 8: tpx_launch<k_xy_fwd<R1, R2, 8>>(nb, 8 * R1, sh, st, xa, Y1, twy); break;
    default: tpx_launch<k_xy_fwd<R1, R2, 16>>(nb, 16 * R1, sh, st, xa, Y1, twy); break;
  }
}
// the packed forward form for x lengths 64, 128 and 256 (R2H = 2, 4, 8)
template <int R2H>
void xy_fwd_rp_launch(unsigned nb, size_t sh, hipStream_t st, const XArgs &xa, int Y1, int Y2, const double2 *twy) {
  switch (Y2) {
    case 4: tpx_launch<k_xy_fwd_rp<R2H, 4>>(nb, 4 * 16, sh, st, xa, Y1, twy); break;
    case 8: tpx_launch<k_xy_fwd_rp<R2H, 8>>(nb, 8 * 16, sh, st, xa, Y1, twy); break;
    default: tpx_launch<k_xy_fwd_rp<R2H, 16>>(nb, 16 * 16, sh, st, xa, Y1, twy); break;
  }
}
template <int R1, int R2, bool Ow, bool Oop>
void xy_inv_launch_t(unsigned nb, size_t sh, hipStream_t st, const XInvArgs &xi, int Y1, int Y2, const double2 *twy) {
  switch (Y2) {
    case 4: tpx_launch<k_xy_inv<R1, R2, 4, Ow, Oop>>(nb, 4 * R1, sh, st, xi, Y1, twy); break;
    case 8: tpx_launch<k_xy_inv<R1, R2, 8, Ow, Oop>>(nb, 8 * R1, sh, st, xi, Y1, twy); break;
    default: tpx_launch<k_xy_inv<R1, R2, 16, Ow, Oop>>(nb, 16 * R1, sh, st, xi, Y1, twy); break;
  }
}
template <int R1, int R2>
void xy_inv_launch(unsigned nb, size_t sh, hipStream_t st, const XInvArgs &xi, int Y1, int Y2, const double2 *twy) {
  if (xi.overwrite) xy_inv_launch_t<R1, R2, true, false>(nb, sh, st, xi, Y1, Y2, twy);
  else if (xi.Pout) xy_inv_launch_t<R1, R2, false, true>(nb, sh, st, xi, Y1, Y2, twy);
  else xy_inv_launch_t<R1, R2, false, false>(nb, sh, st, xi, Y1, Y2, twy);
}
// LDS of a pass-2 block: the run of y1 nz elements, one spare slot every 16,
// then the nz z twiddles and the nz lambda_z (76 KB at nz = 256, y1 = 16: two
// blocks a CU)
size_t yz_shared_bytes(int y1, int nz) {
  return sizeof(double2) * (size_t)y1 * (nz + nz / 16) + (sizeof(double2) + sizeof(double)) * (size_t)nz;
}
template <int Z1, int Z2>
void yz_launch(unsigned nb, size_t sh, hipStream_t st, double2 *spec, int mxh, int Y1, int Y2, const tpx_ctx *c,
               const double *done) {
  switch (Y1) {
    case 4: tpx_launch<k_yz_pass<4, Z1, Z2>>(nb, k_yz_pass<4, Z1, Z2>::kBounds, sh, st, spec, mxh, Y2, c->tw[2], c->lam[0], c->lam[1], c->lam[2], done); break;
    case 8: tpx_launch<k_yz_pass<8, Z1, Z2>>(nb, k_yz_pass<8, Z1, Z2>::kBounds, sh, st, spec, mxh, Y2, c->tw[2], c->lam[0], c->lam[1], c->lam[2], done); break;
    default: tpx_launch<k_yz_pass<16, Z1, Z2>>(nb, k_yz_pass<16, Z1, Z2>::kBounds, sh, st, spec, mxh, Y2, c->tw[2], c->lam[0], c->lam[1], c->lam[2], done); break;
  }
}

void launch_solve_state(tpx_ctx *c, double *dst) {
  tpx_launch<k_solve_state>(1, 64, 0, c->stream, (const double *)c->scal, dst);
  TPX_LAUNCH_CHECK();
}

void launch_solve_begin(tpx_ctx *c) {
  tpx_launch<k_solve_begin>(1, 1, 0, c->stream, c->scal, c->Iteration, c->p.epsilon);
  TPX_LAUNCH_CHECK();
}

void launch_poisson_iteration(tpx_ctx *c, const double *rho_new) {
  launch_poisson_iteration_on(c, rho_new, c->P, c->rhs, c->scal);
}

void launch_residual_coef(tpx_ctx *c, const double *const cf[3], const double *P, const double *rhs, double *res) {
  const Geo &g = c->g;
  tpx_launch<k_residual_coef>((unsigned)((g.n + kThreads - 1) / kThreads), kThreads, 0, c->stream, g, cf[0], cf[1],
                              cf[2], P, rhs, res, (const double *)(c->scal + SC_DONE));
  TPX_LAUNCH_CHECK();
}

// Whether one transform pair fits the three-pass form (every axis 16..256);
// r1, r2 / y1, y2 / z1, z2: the four-step factors of x, y and z.
bool three_pass_shape(const tpx_ctx *c, int &r1, int &r2, int &y1, int &y2, int &z1, int &z2) {
  const Geo &g = c->g;
  return c->fft_passes == 3 && four_step(g.nx, r1, r2) && four_step(g.ny, y1, y2) && four_step(g.nz, z1, z2);
}

// The three passes of one transform pair (passes 1 and 3 as xa, xi ask);
// returns the number of pass-1/3 blocks (the partials pass 3 leaves).
unsigned three_pass_launch(tpx_ctx *c, const XArgs &xa, const XInvArgs &xi, const double *done, int r1, int r2,
                           int y1, int y2) {
  const Geo &g = c->g;
  const int mxh = g.nx / 2 + 1;
  const unsigned nb13 = (unsigned)(g.nz * y1);
  const size_t sh13 = std::max(sizeof(double) * (size_t)y2 * r2 * (r1 + 1), sizeof(double2) * (size_t)mxh * (y2 + 1));
  // pass 1 at x lengths 64, 128 and 256 takes the packed real-row transform
  // unless TPX_XPACK=0 asks for the full-length complex one; pass 3 keeps
  // the full-length inverse (the same half spectrum either way: the packed
  // inverse, whose final stage leaves half the block idle over the P
  // update, measured slower)
  const bool packed = (g.nx == 64 || g.nx == 128 || g.nx == 256) && c->x_packed;
#define L_XYF(A, B) xy_fwd_launch<A, B>(nb13, sh13, c->stream, xa, y1, y2, c->tw[1])
  if (packed && g.nx == 256) xy_fwd_rp_launch<8>(nb13, sh13, c->stream, xa, y1, y2, c->tw[1]);
  else if (packed && g.nx == 128) xy_fwd_rp_launch<4>(nb13, sh13, c->stream, xa, y1, y2, c->tw[1]);
  else if (packed) xy_fwd_rp_launch<2>(nb13, sh13, c->stream, xa, y1, y2, c->tw[1]);
  else { TPX_FOUR_STEP_SWITCH(g.nx, L_XYF) }
#undef L_XYF
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_FFT);
  timing_begin(c, TPX_K_FFT_YZ);
  const unsigned nb2 = (unsigned)(mxh * y2);
  const size_t sh2 = yz_shared_bytes(y1, g.nz);
#define L_YZ(A, B) yz_launch<A, B>(nb2, sh2, c->stream, c->spec[0], mxh, y1, y2, c, done)
  TPX_FOUR_STEP_SWITCH(g.nz, L_YZ)
#undef L_YZ
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_FFT_YZ);
  timing_begin(c, TPX_K_FFT_INV);
#define L_XYI(A, B) xy_inv_launch<A, B>(nb13, sh13, c->stream, xi, y1, y2, c->tw[1])
  TPX_FOUR_STEP_SWITCH(g.nx, L_XYI)
#undef L_XYI
  TPX_LAUNCH_CHECK();
  return nb13;
}

void launch_poisson_iteration_on(tpx_ctx *c, const double *rho_new, double *P, const double *rhs, double *scal,
                                 const double *res, int sum_slot, double *Pout) {
  const Geo &g = c->g;
  int mxh = g.nx / 2 + 1;
  long nlines = (long)g.ny * g.nz;
  int r1, r2;
  timing_begin(c, TPX_K_FFT);
  const double *done = scal + SC_DONE;
  // XArgs.scal serves only the mean of rhs, read at SC_SUM_RHS: point it so
  // that slot is sum_slot
  XArgs xa{g, P, rho_new, rhs, scal + (sum_slot - SC_SUM_RHS), c->spec[0], c->tw[0], (double)g.n, done, res};
  XInvArgs xi{g, P, c->spec[0], c->tw[0], c->partials, (double)g.n, done};
  xi.Pout = Pout;
  const bool x4 = four_step(g.nx, r1, r2);
  int y1, y2, z1, z2;
  // the three-pass form forms the residual itself from rho; a precomputed
  // residual goes through the five-pass kernels
  if (!res && three_pass_shape(c, r1, r2, y1, y2, z1, z2)) {
    const unsigned nb13 = three_pass_launch(c, xa, xi, done, r1, r2, y1, y2);
    tpx_launch<k_finish_iter>(1, 1024, 0, c->stream, c->partials, nb13, scal, c->Iteration, c->p.epsilon);
    TPX_LAUNCH_CHECK();
    timing_end(c, TPX_K_FFT_INV);
    return;
  }
  const int cx = x4 ? 128 / r1 : lines_per_block(g.nx);
  const unsigned nbx = (unsigned)((nlines + cx - 1) / cx);
  const size_t shx4 = x4 ? sizeof(double2) * (size_t)cx * r2 * (r1 + 1) : 0;
  const size_t shx = sizeof(double2) * (size_t)g.nx * cx;
  if (x4) {
#define L_XF(A, B) tpx_launch<k_xfwd4<A, B>>(nbx, cx * A, shx4 / 2, c->stream, xa, cx)
    TPX_FOUR_STEP_SWITCH(g.nx, L_XF)
#undef L_XF
  } else {
    tpx_launch<k_xfwd>(nbx, kThreads, shx, c->stream, xa, cx);
  }
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_FFT);
  timing_begin(c, TPX_K_FFT_YZ);
  const bool y4 = four_step(g.ny, r1, r2);
  if (y4) {
    const int C = 128 / r1;
    const unsigned nb = (unsigned)(((mxh + C - 1) / C) * g.nz);
    const size_t sh = sizeof(double2) * (size_t)C * g.ny;
#define L_YF(A, B) tpx_launch<k_ypass4<A, B, false>>(nb, C * A, sh, c->stream, c->spec[0], mxh, g.ny, g.nz, c->tw[1], C, done)
    TPX_FOUR_STEP_SWITCH(g.ny, L_YF)
#undef L_YF
  } else {
    int ty = lines_per_block(g.ny);
    unsigned nby = (unsigned)(((mxh + ty - 1) / ty) * g.nz);
    tpx_launch<k_ypass<false>>(nby, kThreads, sizeof(double2) * (size_t)ty * g.ny, c->stream, c->spec[0], mxh, g.ny, g.nz,
                                                                                      c->tw[1], ty, done);
  }
  TPX_LAUNCH_CHECK();
  if (four_step(g.nz, r1, r2)) {
    const int C = 128 / r1;
    const unsigned nb = (unsigned)(((mxh + C - 1) / C) * g.ny);
    const size_t sh = sizeof(double2) * (size_t)C * g.nz;
#define L_Z(A, B)                                                                                          \
  tpx_launch<k_zpass4<A, B>>(nb, C * A, sh, c->stream, c->spec[0], mxh, g.ny, g.nz, c->tw[2], c->lam[0], c->lam[1], \
                                               c->lam[2], C, done)
    TPX_FOUR_STEP_SWITCH(g.nz, L_Z)
#undef L_Z
  } else {
    int tz = lines_per_block(g.nz);
    unsigned nbz = (unsigned)(((mxh + tz - 1) / tz) * g.ny);
    tpx_launch<k_zpass>(nbz, kThreads, sizeof(double2) * (size_t)tz * g.nz, c->stream, c->spec[0], mxh, g.ny, g.nz, c->tw[2],
                                                                               c->lam[0], c->lam[1], c->lam[2], tz, done);
  }
  TPX_LAUNCH_CHECK();
  if (y4) {
    four_step(g.ny, r1, r2);
    const int C = 128 / r1;
    const unsigned nb = (unsigned)(((mxh + C - 1) / C) * g.nz);
    const size_t sh = sizeof(double2) * (size_t)C * g.ny;
#define L_YI(A, B) tpx_launch<k_ypass4<A, B, true>>(nb, C * A, sh, c->stream, c->spec[0], mxh, g.ny, g.nz, c->tw[1], C, done)
    TPX_FOUR_STEP_SWITCH(g.ny, L_YI)
#undef L_YI
  } else {
    int ty = lines_per_block(g.ny);
    unsigned nby = (unsigned)(((mxh + ty - 1) / ty) * g.nz);
    tpx_launch<k_ypass<true>>(nby, kThreads, sizeof(double2) * (size_t)ty * g.ny, c->stream, c->spec[0], mxh, g.ny, g.nz,
                                                                                     c->tw[1], ty, done);
  }
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_FFT_YZ);
  timing_begin(c, TPX_K_FFT_INV);
  if (x4) {
#define L_XI(A, B) tpx_launch<k_xinv4<A, B>>(nbx, cx * A, shx4, c->stream, xi, cx)
    TPX_FOUR_STEP_SWITCH(g.nx, L_XI)
#undef L_XI
  } else {
    tpx_launch<k_xinv>(nbx, kThreads, shx, c->stream, xi, cx);
  }
  TPX_LAUNCH_CHECK();
  tpx_launch<k_finish_iter>(1, 1024, 0, c->stream, c->partials, nbx, scal, c->Iteration, c->p.epsilon);
  TPX_LAUNCH_CHECK();
  timing_end(c, TPX_K_FFT_INV);
}

void launch_spectral_apply(tpx_ctx *c, const double *in, double *out, double *scal) {
  const Geo &g = c->g;
  int r1, r2, y1, y2, z1, z2;
  const double *done = scal + SC_DONE;
  if (three_pass_shape(c, r1, r2, y1, y2, z1, z2)) {
    timing_begin(c, TPX_K_FFT);
    XArgs xa{g, nullptr, nullptr, nullptr, scal, c->spec[0], c->tw[0], (double)g.n, done, nullptr};
    xa.in = in;
    XInvArgs xi{g, out, c->spec[0], c->tw[0], c->partials, (double)g.n, done};
    xi.overwrite = 1;
    three_pass_launch(c, xa, xi, done, r1, r2, y1, y2);
    timing_end(c, TPX_K_FFT_INV);
    return;
  }
  // the five-pass form: the input as a precomputed residual, from out = 0;
  // its loop test is undone so that the next application runs
  launch_fill(c, out, 0.0, g.n);
  launch_poisson_iteration_on(c, nullptr, out, nullptr, scal, in);
  launch_fill(c, scal + SC_ITERS, 0.0, 1);
  launch_fill(c, scal + SC_DONE, 0.0, 1);
}

}  // namespace tpx
