// fields.cpp -- device-field handles, particles, file I/O and the Poisson
// facade of the host mirror.
#include "fields.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

void tpx_check(tpx_ctx *ctx, int rc, const char *what) {
  if (rc == TPX_OK) return;
  std::ostringstream m;
  m << what << " failed (" << rc << "): " << (ctx ? tpx_last_error(ctx) : tpx_create_error());
  throw std::runtime_error(m.str());
}

// ---------------------------------------------------------------- tensor0
void tensor0::bind(tpx_ctx *const *ctx, gridsize *size, int field, int comp) {
  ctx_ = ctx;
  size_ = size;
  field_ = field;
  comp_ = comp;
}

size_t tensor0::ghosted_size() const {
  size_t b = 2 * (size_t)size_->bs();
  return ((size_t)size_->Nx() + b) * ((size_t)size_->Ny() + b) * ((size_t)size_->Nz() + b);
}

double tensor0::sum() const {
  double v = 0;
  tpx_check(ctx(), tpx_field_sum(ctx(), field_, comp_, &v), "tensor0::sum");
  return v;
}

double tensor0::mean() const { return sum() / (double)size_->size_tot(); }

double tensor0::max_abs() const {
  double v = 0;
  tpx_check(ctx(), tpx_field_max_abs(ctx(), field_, comp_, &v), "tensor0::max_abs");
  return v;
}

std::vector<double> tensor0::host() const {
  std::vector<double> h(ghosted_size());
  download(h.data());
  return h;
}

void tensor0::download(double *ghosted) const {
  tpx_check(ctx(), tpx_download_field(ctx(), field_, comp_, ghosted, size_->bs()), "tensor0 download");
}

void tensor0::upload(const double *ghosted) {
  tpx_check(ctx(), tpx_upload_field(ctx(), field_, comp_, ghosted, size_->bs()), "tensor0 upload");
}

void tensor0::kill_strong_modes() {
  throw std::runtime_error("kill_strong_modes (elongated box) is outside this library");
}

// ---------------------------------------------------------------- tensor1
void tensor1::bind(tpx_ctx *const *ctx, gridsize *size, int field) {
  x.bind(ctx, size, field, 0);
  y.bind(ctx, size, field, 1);
  z.bind(ctx, size, field, 2);
}

void tensor1::make_mean_U0(double *a) {
  if (x.field() != TPX_RU_NP1)
    throw std::runtime_error("make_mean_U0 is implemented on the device for RU_np1 only");
  tpx_check(x.ctx(), tpx_make_mean_u0(x.ctx(), a), "make_mean_U0");
}

// ---------------------------------------------------------------- files
namespace {

template <class T>
bool read_raw(const std::string &name, std::vector<T> &out, size_t n, size_t offset_elems = 0) {
  std::ifstream f(name, std::ios::binary);
  if (!f) return false;
  out.resize(n);
  f.seekg((std::streamoff)(offset_elems * sizeof(T)));
  f.read(reinterpret_cast<char *>(out.data()), (std::streamsize)(n * sizeof(T)));
  return (size_t)f.gcount() == n * sizeof(T);
}

template <class T>
void write_raw(const std::string &name, const T *p, size_t n, bool append) {
  std::ofstream f(name, std::ios::binary | (append ? std::ios::app : std::ios::trunc));
  if (!f) throw std::runtime_error("cannot write " + name);
  f.write(reinterpret_cast<const char *>(p), (std::streamsize)(n * sizeof(T)));
}

// global C-order interior <-> ghosted local layout
void interior_to_ghosted(const gridsize &g, const double *in, double *out) {
  int b = g.bs();
  long gx = g.Nx() + 2 * b, gy = g.Ny() + 2 * b;
  for (int k = 0; k < g.Nz(); ++k)
    for (int j = 0; j < g.Ny(); ++j)
      for (int i = 0; i < g.Nx(); ++i)
        out[((long)(k + b) * gy + (j + b)) * gx + (i + b)] = in[((long)k * g.Ny() + j) * g.Nx() + i];
}

void ghosted_to_interior(const gridsize &g, const double *in, double *out) {
  int b = g.bs();
  long gx = g.Nx() + 2 * b, gy = g.Ny() + 2 * b;
  for (int k = 0; k < g.Nz(); ++k)
    for (int j = 0; j < g.Ny(); ++j)
      for (int i = 0; i < g.Nx(); ++i)
        out[((long)k * g.Ny() + j) * g.Nx() + i] = in[((long)(k + b) * gy + (j + b)) * gx + (i + b)];
}

void read_component(gridsize &g, tensor0 &T, const char *adrs, int comp) {
  size_t n = (size_t)g.size_tot();
  std::vector<double> in;
  if (!read_raw(adrs, in, n, (size_t)comp * n)) throw std::runtime_error(std::string("cannot read ") + adrs);
  std::vector<double> gh(T.ghosted_size(), 0.0);
  interior_to_ghosted(g, in.data(), gh.data());
  T.upload(gh.data());
}

void write_component(gridsize &g, tensor0 &T, const char *adrs, bool append) {
  std::vector<double> gh = T.host();
  std::vector<double> out((size_t)g.size_tot());
  ghosted_to_interior(g, gh.data(), out.data());
  write_raw(adrs, out.data(), out.size(), append);
}

}  // namespace

void communicator::read(tensor0 &T, const char *adrs) { read_component(*size_, T, adrs, 0); }

void communicator::read(tensor1 &T, const char *adrs) {
  read_component(*size_, T.x, adrs, 0);
  read_component(*size_, T.y, adrs, 1);
  read_component(*size_, T.z, adrs, 2);
}

std::vector<double> communicator::read_host(const char *adrs, int comp) {
  size_t n = (size_t)size_->size_tot();
  std::vector<double> in;
  if (!read_raw(adrs, in, n, (size_t)comp * n)) throw std::runtime_error(std::string("cannot read ") + adrs);
  std::vector<double> gh(ghosted_count(), 0.0);
  interior_to_ghosted(*size_, in.data(), gh.data());
  return gh;
}

void communicator::write_host(const std::vector<double> &ghosted, const char *adrs, bool append) {
  std::vector<double> out((size_t)size_->size_tot());
  ghosted_to_interior(*size_, ghosted.data(), out.data());
  write_raw(adrs, out.data(), out.size(), append);
}

size_t communicator::ghosted_count() const {
  size_t b = 2 * (size_t)size_->bs();
  return ((size_t)size_->Nx() + b) * ((size_t)size_->Ny() + b) * ((size_t)size_->Nz() + b);
}

void communicator::write(tensor0 &T, const char *adrs) { write_component(*size_, T, adrs, false); }

void communicator::write(tensor1 &T, const char *adrs) {
  write_component(*size_, T.x, adrs, false);
  write_component(*size_, T.y, adrs, true);
  write_component(*size_, T.z, adrs, true);
}

// ---------------------------------------------------------------- particles
void particle::Load_All() {
  Np = param_->N0();
  const char *names[7] = {"x", "y", "z", "u", "v", "w", "T"};
  std::vector<double> *dst[7] = {&x, &y, &z, &u, &v, &w, &T};
  for (int q = 0; q < 7; ++q) {
    std::string f = std::string("Restart_Part_") + names[q] + ".bin";
    if (!read_raw(f, *dst[q], (size_t)Np)) throw std::runtime_error("cannot read " + f);
  }
  if (!read_raw(std::string("Restart_Part_flag.bin"), flag, (size_t)Np))
    throw std::runtime_error("cannot read Restart_Part_flag.bin");
  std::vector<double> tau, eps;
  bool has_tau = read_raw(std::string("Restart_Part_tau.bin"), tau, (size_t)Np);
  bool has_eps = read_raw(std::string("Restart_Part_eps.bin"), eps, (size_t)Np);
  tpx_check(*ctx_, tpx_upload_particles(*ctx_, Np, x.data(), y.data(), z.data(), u.data(), v.data(), w.data(),
                                        T.data(), flag.data(), has_tau ? tau.data() : nullptr,
                                        has_eps ? eps.data() : nullptr),
            "particle::Load_All");
}

void particle::load_random(double Lx, double Ly, double Lz) {
  Np = param_->N0();
  x.clear(); y.clear(); z.clear(); u.clear(); v.clear(); w.clear(); T.clear(); flag.clear();
  for (int i = 0; i < Np; ++i) {
    x.push_back(((double)rand() / (double)RAND_MAX) * Lx);
    u.push_back(0);
    y.push_back(((double)rand() / (double)RAND_MAX) * Ly);
    v.push_back(0);
    z.push_back(((double)rand() / (double)RAND_MAX) * Lz);
    w.push_back(0);
    T.push_back(param_->T0());
    flag.push_back(i);
  }
  tpx_check(*ctx_, tpx_upload_particles(*ctx_, Np, x.data(), y.data(), z.data(), u.data(), v.data(), w.data(),
                                        T.data(), flag.data(), nullptr, nullptr),
            "particle::load_random");
}

void particle::sync() {
  int64_t n = 0;
  tpx_check(*ctx_, tpx_num_particles(*ctx_, &n), "particle count");
  Np = (int)n;
  for (auto *v : {&x, &y, &z, &u, &v, &w, &T}) v->resize((size_t)n);
  flag.resize((size_t)n);
  tpx_check(*ctx_, tpx_download_particles(*ctx_, TPX_BASE, x.data(), y.data(), z.data(), u.data(), v.data(),
                                          w.data(), T.data(), flag.data()),
            "particle download");
}

void particle::Store_All(int num_ts) {
  sync();
  const char *names[7] = {"x", "y", "z", "u", "v", "w", "T"};
  std::vector<double> *src[7] = {&x, &y, &z, &u, &v, &w, &T};
  for (int q = 0; q < 7; ++q) {
    std::ostringstream f;
    f << param_->data_dir() << "Part_" << names[q] << "_" << num_ts << ".bin";
    write_raw(f.str(), src[q]->data(), src[q]->size(), false);
  }
  std::ostringstream f;
  f << param_->data_dir() << "Part_flag_" << num_ts << ".bin";
  write_raw(f.str(), flag.data(), flag.size(), false);
}

// ---------------------------------------------------------------- poisson
void poisson::Solve(tensor1 &C, tensor0 &P, tensor0 &RHS) {
  // the grid's own solve (grid.cpp:851-855): C = 1/Rho_face is formed on the
  // device from Rho_new, so grid::Solve_Poisson passes an unbound C
  if (!C.x.ctx()) {
    if (P.field() != TPX_P || RHS.field() != TPX_RHS_POIS)
      throw std::runtime_error("poisson::Solve without a coefficient runs on the grid's own P and RHS_Pois");
    tpx_check(P.ctx(), tpx_solve_poisson(P.ctx(), 0, &num_iteration_), "poisson::Solve");
    return;
  }
  // any other coefficient, pressure and right-hand side (poisson.cpp:115-117)
  std::vector<double> cx = C.x.host(), cy = C.y.host(), cz = C.z.host(), p = P.host(), r = RHS.host();
  tpx_check(P.ctx(), tpx_poisson_solve(P.ctx(), cx.data(), cy.data(), cz.data(), p.data(), r.data(),
                                       P.size()->bs(), &num_iteration_),
            "poisson::Solve");
  P.upload(p.data());
}
