/*
 * or_fft.c -- radix-2 complex FFT for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Stands in for Sandia fft_3d + FFTW 2.1.5 (external, not in the reference
 * tree; call sites poisson.cpp:87,431,453).  Same contract as those calls:
 * in-place, unnormalised, x fastest, complex interleaved {re,im}.  Pinned in
 * tests against a brute-force DFT and against numpy.fft fixtures.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static int ilog2i(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }

/* Twiddle table w[m] = exp(sign * 2 pi i m / n), m < n/2. */
static double *make_twiddles(int n, int sign)
{
  double *w = (double *)malloc(sizeof(double) * (size_t)(n > 1 ? n : 2));
  for (int m = 0; m < n / 2; ++m) {
    double a = 2.0 * 3.141592653589793 * (double)m / (double)n;
    w[2 * m] = cos(a);
    w[2 * m + 1] = (double)sign * sin(a);
  }
  return w;
}

/* Iterative decimation-in-time radix-2 transform of one contiguous line. */
static void fft_line(double *a, int n, const double *w)
{
  int lg = ilog2i(n);
  for (int i = 0; i < n; ++i) {           /* bit reversal */
    int r = 0;
    for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
    if (r > i) {
      double tr = a[2 * i], ti = a[2 * i + 1];
      a[2 * i] = a[2 * r]; a[2 * i + 1] = a[2 * r + 1];
      a[2 * r] = tr; a[2 * r + 1] = ti;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, step = n / len;
    for (int base = 0; base < n; base += len) {
      for (int m = 0; m < half; ++m) {
        double wr = w[2 * m * step], wi = w[2 * m * step + 1];
        double *p = a + 2 * (base + m), *q = a + 2 * (base + m + half);
        double vr = q[0] * wr - q[1] * wi;
        double vi = q[0] * wi + q[1] * wr;
        double ur = p[0], ui = p[1];
        p[0] = ur + vr; p[1] = ui + vi;
        q[0] = ur - vr; q[1] = ui - vi;
      }
    }
  }
}

void or_fft3d(double *data, int nx, int ny, int nz, int sign)
{
  double *wx = make_twiddles(nx, sign), *wy = make_twiddles(ny, sign),
         *wz = make_twiddles(nz, sign);
  long nxy = (long)nx * ny;
  /* x lines are contiguous */
#pragma omp parallel for schedule(static)
  for (long line = 0; line < (long)ny * nz; ++line)
    fft_line(data + 2 * line * nx, nx, wx);
  /* y lines: stride nx */
#pragma omp parallel
  {
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)ny);
#pragma omp for schedule(static)
    for (long line = 0; line < (long)nx * nz; ++line) {
      long k = line / nx, i = line % nx;
      double *base = data + 2 * (k * nxy + i);
      for (int j = 0; j < ny; ++j) { tmp[2 * j] = base[2 * (long)j * nx]; tmp[2 * j + 1] = base[2 * (long)j * nx + 1]; }
      fft_line(tmp, ny, wy);
      for (int j = 0; j < ny; ++j) { base[2 * (long)j * nx] = tmp[2 * j]; base[2 * (long)j * nx + 1] = tmp[2 * j + 1]; }
    }
    free(tmp);
  }
  /* z lines: stride nx*ny */
#pragma omp parallel
  {
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)nz);
#pragma omp for schedule(static)
    for (long line = 0; line < nxy; ++line) {
      double *base = data + 2 * line;
      for (int k = 0; k < nz; ++k) { tmp[2 * k] = base[2 * (long)k * nxy]; tmp[2 * k + 1] = base[2 * (long)k * nxy + 1]; }
      fft_line(tmp, nz, wz);
      for (int k = 0; k < nz; ++k) { base[2 * (long)k * nxy] = tmp[2 * k]; base[2 * (long)k * nxy + 1] = tmp[2 * k + 1]; }
    }
    free(tmp);
  }
  free(wx); free(wy); free(wz);
}
