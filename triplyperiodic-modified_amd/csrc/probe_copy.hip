// probe_copy.hip -- streaming-bandwidth probe of the device (gfx950): what a
// plain hand-written copy, read and write reach, the practical ceiling the
// step's kernels are compared with, and the calibration kernel of the
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters (its bytes are known exactly).
// A measurement tool of the benchmark, not part of the step's ABI.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kThreads = 256;

// 16-byte accesses, each thread striding over the array by the grid's width
__global__ void __launch_bounds__(kThreads) k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long n2) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n2; i += (long)gridDim.x * kThreads) dst[i] = src[i];
}

// every element read; the per-thread sums are folded into one value per
// block so the loads cannot be dropped
__global__ void __launch_bounds__(kThreads) k_read(const double2 *__restrict__ src, double *__restrict__ out, long n2) {
  double s = 0.0;
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n2; i += (long)gridDim.x * kThreads) {
    double2 v = src[i];
    s += v.x + v.y;
  }
  __shared__ double red[kThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kThreads) k_write(double2 *__restrict__ dst, long n2, double v) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n2; i += (long)gridDim.x * kThreads)
    dst[i] = make_double2(v, v);
}

}  // namespace

extern "C" {

// Copies, reads and writes `bytes` bytes (a multiple of 16) `reps` times each
// with `blocks` workgroups of 256 threads (0: 32 per CU); out[0..2] receive
// the best GB/s of the copy (bytes read + bytes written), the read and the
// write.  Returns 0, or -1 on a bad argument, -2 on a HIP error.
int tpx_probe_bandwidth(int64_t bytes, int reps, int blocks, double *out) {
  if (bytes <= 0 || bytes % 16 != 0 || reps <= 0 || !out) return -1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -2;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -2;
  const unsigned nb = blocks > 0 ? (unsigned)blocks : (unsigned)(32 * cus);
  const long n2 = (long)(bytes / 16);
  double2 *a = nullptr, *b = nullptr;
  double *part = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
      hipMalloc(&part, sizeof(double) * nb) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = -2;
  } else {
    hipLaunchKernelGGL(k_write, dim3(nb), dim3(kThreads), 0, 0, a, n2, 1.0);
    hipLaunchKernelGGL(k_write, dim3(nb), dim3(kThreads), 0, 0, b, n2, 0.0);
    double best[3] = {0, 0, 0};
    for (int kind = 0; kind < 3 && rc == 0; ++kind) {
      for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, 0);
        if (kind == 0) hipLaunchKernelGGL(k_copy, dim3(nb), dim3(kThreads), 0, 0, a, b, n2);
        if (kind == 1) hipLaunchKernelGGL(k_read, dim3(nb), dim3(kThreads), 0, 0, a, part, n2);
        if (kind == 2) hipLaunchKernelGGL(k_write, dim3(nb), dim3(kThreads), 0, 0, b, n2, 2.0);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = -2; break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double moved = (kind == 0 ? 2.0 : 1.0) * (double)bytes;
        const double gbps = moved / (ms * 1e-3) / 1e9;
        if (gbps > best[kind]) best[kind] = gbps;
      }
    }
    if (hipDeviceSynchronize() != hipSuccess) rc = -2;
    for (int k = 0; k < 3; ++k) out[k] = best[k];
  }
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (part) (void)hipFree(part);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

}  // extern "C"
