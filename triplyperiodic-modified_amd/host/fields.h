// fields.h -- host handles with the reference's names for the device-resident
// fields (tensor0.h:6-104, tensor1.h:6-85), the particle set (particle.h),
// the Poisson solver (poisson.h:79) and file I/O (communicator.h).
//
// A tensor0 here owns no memory: it names one field (and component) of a
// device context, and its operations forward to the C ABI.  Host copies are
// taken explicitly in the reference's ghosted layout (Nx+2bs)(Ny+2bs)(Nz+2bs),
// x fastest, so code that inspects values keeps its indexing.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tpx.h"
#include "params.h"

// Throws with the context's last error when an ABI call fails.
void tpx_check(tpx_ctx *ctx, int rc, const char *what);

class tensor0 {
  tpx_ctx *const *ctx_ = nullptr;   // the owning grid's context slot
  gridsize *size_ = nullptr;
  int field_ = -1, comp_ = 0;

 public:
  tensor0() {}
  void bind(tpx_ctx *const *ctx, gridsize *size, int field, int comp);
  int field() const { return field_; }
  int comp() const { return comp_; }
  tpx_ctx *ctx() const { return ctx_ ? *ctx_ : nullptr; }
  gridsize *size() const { return size_; }
  // number of doubles of a ghosted local array
  size_t ghosted_size() const;
  // reductions over the interior (tensor0.cpp:364-396, 466-478)
  double sum() const;
  double mean() const;
  double max_abs() const;
  // host copies, ghosted layout with periodic ghosts
  std::vector<double> host() const;
  void download(double *ghosted) const;
  void upload(const double *ghosted);
  // elongated-box mode filter (tensor0.cpp) is outside this library
  void kill_strong_modes();
};

class tensor1 {
 public:
  tensor0 x, y, z;
  void bind(tpx_ctx *const *ctx, gridsize *size, int field);
  // RU_np1 <- RU_np1 - mean + a, per component (tensor1.cpp:406-411); the
  // device implements it for RU_np1, where the driver calls it.
  void make_mean_U0(double *a);
  double sum() const { return x.sum() + y.sum() + z.sum(); }
};

// Particles (particle.h:61-152).  The state lives on the device; the vectors
// below are filled by sync() with the current base state, in the device's
// order (identify particles by flag).
class particle {
  tpx_ctx *const *ctx_ = nullptr;
  params *param_ = nullptr;

 public:
  int Np = 0;
  std::vector<double> x, y, z, u, v, w, T;
  std::vector<int> flag;
  void bind(tpx_ctx *const *ctx, params *p) { ctx_ = ctx; param_ = p; }
  // reads Restart_Part_{x,y,z,u,v,w,T,flag}.bin (N0 particles) from the
  // working directory (particle.cpp:899-963) and uploads them; optional
  // Restart_Part_tau.bin / Restart_Part_eps.bin carry per-particle response
  // times and absorption efficiencies
  void Load_All();
  // N0 particles at rest at positions from the C library's rand(), one
  // draw per coordinate, as particle::load_random (particle.cpp:903-935)
  void load_random(double Lx, double Ly, double Lz);
  // copies the base state back into the vectors
  void sync();
  // writes Part_{x,...,flag}_<n>.bin into Data_dir (particle.cpp:848-897)
  void Store_All(int num_ts);
};

// Raw FP64 global arrays, C order, tensor1 as three concatenated components
// (communicator.cpp:65-106).
class communicator {
  gridsize *size_;

 public:
  communicator(gridsize *size, params *, proc *) : size_(size) {}
  void read(tensor0 &T, const char *adrs);
  void read(tensor1 &T, const char *adrs);
  void write(tensor0 &T, const char *adrs);
  void write(tensor1 &T, const char *adrs);
  // the same files for fields kept on the host in the ghosted layout: one
  // component of a file, and one array written (or appended, for the later
  // components of a tensor1)
  std::vector<double> read_host(const char *adrs, int comp = 0);
  void write_host(const std::vector<double> &ghosted, const char *adrs, bool append = false);
  size_t ghosted_count() const;
};

class poisson {
  int num_iteration_ = 0;

 public:
  // Solves div(C grad P) = RHS (poisson.cpp:115-117, 391-462).  With an
  // unbound C it is the grid's own solve (P and RHS_Pois, C = 1/Rho_face from
  // Rho_new on the device); otherwise C, P and RHS are taken as they are.
  void Solve(tensor1 &C, tensor0 &P, tensor0 &RHS);
  int num_iteration() const { return num_iteration_; }
};
