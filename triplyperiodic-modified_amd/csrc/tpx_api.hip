// tpx_api.hip -- the C ABI (include/tpx.h): device state, buffer rotation
// and the order of kernel launches for each reference method and for the
// fused step.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <initializer_list>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <string>
#include <vector>

#include "tpx_internal.h"

using namespace tpx;

namespace tpx {
thread_local Recorder *g_rec = nullptr;
}

// the static scratch of tpx_poisson_solve_device (see static_scratch below)
constexpr long kStaticScratchDoubles = 1L << 21;
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

namespace {

thread_local std::string g_create_error;

// The scratch of tpx_poisson_solve_device: one file-scope device array per
// GPU, holding the residual of one coefficient solve (2^21 doubles, 16 MiB:
// grids of up to 128^3 cells).  Host code never names the array itself (its
// name is a host-side stand-in, not a device address): tpx_create looks up
// its address on the current device once with hipGetSymbolAddress and every
// use goes through that pointer.  The array lives until the process ends.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

void hip_check_symbol(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
}

// the array's address on device `dev` (current), looked up on first use
double *static_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the static scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    hip_check_symbol(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  return g_scratch_ptr[dev];
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

template <class F>
int guarded(tpx_ctx *c, F &&f) {
  if (!c) return TPX_E_ARG;
  try {
    return f();
  } catch (const std::invalid_argument &e) {
    c->err = e.what();
    return TPX_E_ARG;
  } catch (const std::logic_error &e) {
    c->err = e.what();
    return TPX_E_STATE;
  } catch (const std::bad_alloc &e) {
    c->err = "out of memory";
    return TPX_E_OOM;
  } catch (const std::exception &e) {
    c->err = e.what();
    return TPX_E_HIP;
  }
}

// coefficient pack for substep rk (grid.cpp:18-25 tables and params)
Coef coef(const tpx_ctx *c, int rk) {
  Coef k{};
  k.dt = c->p.dt;
  k.pre = c->pre[rk];
  k.post = c->post[rk];
  k.a_drag = -c->mp / c->Vcell;
  k.b_visc = c->p.mu0 / 3.;
  k.c_visc = c->p.mu0;
  k.A = c->A;
  k.rho0 = c->p.rho0;
  for (int d = 0; d < 3; ++d) k.g[d] = c->gv[d];
  k.k_over_cp = c->p.k / c->p.cp;
  k.R = c->R;
  k.cp = c->p.cp;
  k.cv = c->p.cv;
  k.Vcell = c->Vcell;
  k.rk = rk;
  k.cooling = c->p.cooling;
  return k;
}

// ---- buffer bookkeeping -------------------------------------------------
int pick_free(std::initializer_list<int> used, int nbuf) {
  for (int b = 0; b < nbuf; ++b)
    if (std::find(used.begin(), used.end(), b) == used.end()) return b;
  throw std::logic_error("no free field buffer");
}

int ru_np1_read(const tpx_ctx *c) { return c->ru_np1_is_base ? c->ru_base : c->ru_np1; }
int rho_np1_read(const tpx_ctx *c) { return c->rho_np1_is_base ? c->rho_base : c->rho_np1; }

// make RU_np1 / RU_new writable buffers for substep rk
void prepare_ru_outputs(tpx_ctx *c, int rk) {
  if (c->ru_np1_is_base) {
    c->ru_np1 = pick_free({c->ru_base, c->ru_int, c->ru_new}, 5);
    c->ru_np1_is_base = false;
    c->ru_new = -1;
  }
  if (rk == 3) c->ru_new = c->ru_np1;
  else if (c->ru_new < 0 || c->ru_new == c->ru_int || c->ru_new == c->ru_base || c->ru_new == c->ru_np1)
    c->ru_new = pick_free({c->ru_base, c->ru_int, c->ru_np1}, 5);
}

void prepare_rho_outputs(tpx_ctx *c, int rk) {
  if (c->rho_np1_is_base) {
    c->rho_np1 = pick_free({c->rho_base, c->rho_int, c->rho_new}, 5);
    c->rho_np1_is_base = false;
    c->rho_new = -1;
  }
  if (rk == 3) c->rho_new = c->rho_np1;
  else if (c->rho_new < 0 || c->rho_new == c->rho_int || c->rho_new == c->rho_base || c->rho_new == c->rho_np1)
    c->rho_new = pick_free({c->rho_base, c->rho_int, c->rho_np1}, 5);
}

void ru_ptrs(tpx_ctx *c, int b, const double *out[3]) {
  for (int d = 0; d < 3; ++d) out[d] = c->ru[b][d];
}
void ru_wptrs(tpx_ctx *c, int b, double *out[3]) {
  for (int d = 0; d < 3; ++d) out[d] = c->ru[b][d];
}

// the passive scalar's four buffers, allocated (zero) on first use
void ensure_passive_scalar(tpx_ctx *c) {
  if (c->ps[0]) return;
  for (auto &f : c->ps) {
    if (hipMalloc(&f, sizeof(double) * (size_t)c->g.n) != hipSuccess) throw std::bad_alloc();
    hip_check(hipMemsetAsync(f, 0, sizeof(double) * (size_t)c->g.n, c->stream), "hipMemsetAsync");
  }
}

// the passive vector's buffers (RV's four sets, Q, the raw right-hand side of
// Q's solve, the type-1 source's scratch: zero) and the constant density
// Rho_forV as a field, allocated on first use
void ensure_passive_vector(tpx_ctx *c) {
  if (c->Qv) return;
  const size_t bytes = sizeof(double) * (size_t)c->g.n;
  auto zero = [&](double *&f) {
    if (hipMalloc(&f, bytes) != hipSuccess) throw std::bad_alloc();
    hip_check(hipMemsetAsync(f, 0, bytes, c->stream), "hipMemsetAsync");
  };
  for (auto &set : c->rv)
    for (auto &f : set) zero(f);
  zero(c->rhs_q);
  zero(c->rv_filt);
  if (hipMalloc(&c->rho_forv_field, bytes) != hipSuccess) throw std::bad_alloc();
  launch_fill(c, c->rho_forv_field, c->p.rho_forv, c->g.n);
  zero(c->Qv);
}

double *field_buffer(tpx_ctx *c, int field, int comp) {
  if (field >= TPX_PASSIVE_SCALAR && field <= TPX_PASSIVE_SCALAR_NP1) ensure_passive_scalar(c);
  if ((field >= TPX_RV && field <= TPX_RV_NP1) || field == TPX_Q || field == TPX_RHS_POIS_Q) ensure_passive_vector(c);
  switch (field) {
  case TPX_RU: return c->ru[c->ru_base][comp];
  case TPX_RU_INT: return c->ru[c->ru_int][comp];
  case TPX_RU_NEW: return c->ru_new >= 0 ? c->ru[c->ru_new][comp] : nullptr;
  case TPX_RU_NP1: return c->ru[ru_np1_read(c)][comp];
  case TPX_RHS_RU: return c->S[comp];
  case TPX_P: return c->P;
  case TPX_RHS_POIS: return c->rhs;
  case TPX_C: return c->C;
  case TPX_RHO: return c->rho[c->rho_base];
  case TPX_RHO_INT: return c->rho[c->rho_int];
  case TPX_RHO_NEW: return c->rho_new >= 0 ? c->rho[c->rho_new] : nullptr;
  case TPX_RHO_NP1: return c->rho[rho_np1_read(c)];
  case TPX_DIVERGENCE: return c->div_valid ? c->divergence : nullptr;
  case TPX_RHS_PART_TEMP: return c->Q;
  case TPX_PASSIVE_SCALAR: return c->ps[c->ps_base];
  case TPX_PASSIVE_SCALAR_INT: return c->ps[c->ps_int];
  case TPX_PASSIVE_SCALAR_NEW: return c->ps[c->ps_new];
  case TPX_PASSIVE_SCALAR_NP1: return c->ps[c->ps_np1];
  case TPX_RV: return c->rv[c->rv_base][comp];
  case TPX_RV_INT: return c->rv[c->rv_int][comp];
  case TPX_RV_NEW: return c->rv[c->rv_new][comp];
  case TPX_RV_NP1: return c->rv[c->rv_np1][comp];
  case TPX_Q: return c->Qv;
  case TPX_RHS_POIS_Q: return c->rhs_q;
  default: return nullptr;
  }
}

bool is_vector(int field) { return field >= TPX_RU && field <= TPX_RV_NP1; }

void sync(tpx_ctx *c) { hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize"); }

double read_scalar(tpx_ctx *c, int slot) {
  hip_check(hipMemcpyAsync(c->h_scal + slot, c->scal + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpyAsync");
  sync(c);
  return c->h_scal[slot];
}

void write_scalar(tpx_ctx *c, int slot, double v) {
  c->h_scal[slot] = v;
  hip_check(hipMemcpyAsync(c->scal + slot, c->h_scal + slot, sizeof(double), hipMemcpyHostToDevice, c->stream),
            "hipMemcpyAsync");
  sync(c);
}

bool heat_active(const tpx_ctx *c) { return c->p.nu * TPX_PI * c->p.dp * c->p.k != 0.0; }

// ---- method bodies ------------------------------------------------------
void do_update_rho(tpx_ctx *c, int rk) {
  Coef k = coef(c, rk);
  int np1_in = rho_np1_read(c);
  prepare_rho_outputs(c, rk);
  const double *ri[3];
  ru_ptrs(c, c->ru_int, ri);
  launch_update_rho(c, k, ri, c->rho[c->rho_base], c->rho[np1_in], c->rho[c->rho_np1], c->rho[c->rho_new]);
}

void zero_particle_fields(tpx_ctx *c) {
  for (int d = 0; d < 3; ++d) launch_fill(c, c->S[d], 0.0, c->g.n);
  launch_fill(c, c->Q, 0.0, c->g.n);
  launch_fill(c, c->scal + SC_SUM_Q, 0.0, 1);
  c->q_zero = true;
}

// Update_Particle (grid.cpp:865-888) as one method: includes the
// statistics-only concentration deposit at x_int.
void do_update_particle(tpx_ctx *c, int rk) {
  Coef k = coef(c, rk);
  Particles &pt = c->pt;
  if (pt.n == 0) {
    zero_particle_fields(c);
    return;
  }
  const double *ri[3];
  ru_ptrs(c, c->ru_int, ri);
  if (c->stats_on && rk == 3) {
    stats_capture_particles(c, ri, c->rho[c->rho_int]);
  } else {
    // the bins of x_int are current only when the particles are re-binned
    // every substep (two-way coupling or heat exchange)
    if (!(c->p.twc || heat_active(c))) launch_sort_stage(c, TPX_INT, true);
    launch_concentration(c);
  }
  launch_particle_advance(c, k, ri, c->rho[c->rho_int], true);
  if (c->p.twc) launch_velocity_deposit(c);
  else for (int d = 0; d < 3; ++d) launch_fill(c, c->S[d], 0.0, c->g.n);
  launch_sort_new(c);
  c->binned_this_step = true;
  if (heat_active(c)) {
    launch_heat_exchange(c, TPX_NEW, c->rho[c->rho_new], SC_P0_NEW);
    c->q_zero = false;
  } else {
    launch_fill(c, pt.rhs_t, 0.0, pt.n);
    if (!c->q_zero) {
      launch_fill(c, c->Q, 0.0, c->g.n);
      launch_fill(c, c->scal + SC_SUM_Q, 0.0, 1);
      c->q_zero = true;
    }
  }
}

// grid::Update_Passive_Scalar (grid.cpp:711-745)
void do_update_passive_scalar(tpx_ctx *c, int rk) {
  ensure_passive_scalar(c);
  Coef k = coef(c, rk);
  const double *ri[3];
  ru_ptrs(c, c->ru_int, ri);
  launch_passive_scalar(c, k, ri, c->rho[c->rho_int]);
}

void do_update_ru_wop(tpx_ctx *c, int rk, bool with_rho) {
  Coef k = coef(c, rk);
  int np1_in = ru_np1_read(c);
  int rho_np1_in = rho_np1_read(c);
  prepare_ru_outputs(c, rk);
  if (with_rho) prepare_rho_outputs(c, rk);
  const double *ri[3], *rb[3], *rn[3];
  double *on[3], *ow[3];
  ru_ptrs(c, c->ru_int, ri);
  ru_ptrs(c, c->ru_base, rb);
  ru_ptrs(c, np1_in, rn);
  ru_wptrs(c, c->ru_np1, on);
  ru_wptrs(c, c->ru_new, ow);
  bool has_s = c->p.twc && c->pt.n > 0;
  const double *S[3] = {c->S[0], c->S[1], c->S[2]};
  launch_rhs_ru(c, k, with_rho, ri, c->rho[c->rho_int], rb, rn, has_s ? S : nullptr, on, ow,
                c->rho[c->rho_base], c->rho[rho_np1_in],
                with_rho ? c->rho[c->rho_np1] : nullptr, with_rho ? c->rho[c->rho_new] : nullptr);
}

// Solve_Poisson's loop (poisson.cpp:391-462).  The loop test runs on the
// device after every iteration, and iterations queued after it fails skip
// themselves, so the host queues iterations in batches and waits once per
// batch instead of once per iteration: first as many as the last solve of
// this substep took, then two at a time.
// The host side of every Poisson solve: the device counts the iterations and
// evaluates the stopping rule (poisson.cpp:397) itself, so iterations are
// queued in batches (the first as long as `guess`, then two at a time) and
// the host waits once per batch.  Returns the iterations that ran.
int solve_loop(tpx_ctx *c, int guess, const std::function<void(int)> &iterate) {
  launch_solve_begin(c);
  int queued = 0, batch = std::max(1, guess);
  double st[2] = {0.0, 1.0};  // SC_ITERS, SC_DONE
  while (queued < c->Iteration) {
    const int nb = std::min(batch, c->Iteration - queued);
    for (int b = 0; b < nb; ++b) iterate(queued + b);
    queued += nb;
    hip_check(hipMemcpyAsync(c->h_scal + SC_ITERS, c->scal + SC_ITERS, 2 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream),
              "hipMemcpyAsync");
    sync(c);
    st[0] = c->h_scal[SC_ITERS];
    st[1] = c->h_scal[SC_DONE];
    if (st[1] != 0.0) break;
    batch = 2;
  }
  return (int)st[0];
}

void do_solve_poisson(tpx_ctx *c, int rk, int *iters, const std::function<void(int)> &before_iteration = {}) {
  const int i = solve_loop(c, c->iter_guess[rk & 3], [&](int it) {
    if (before_iteration) before_iteration(it);
    launch_poisson_iteration(c, c->rho[c->rho_new]);
  });
  c->iter_guess[rk & 3] = i;
  c->num_iteration = i;
  c->total_iterations += i;
  if (iters) *iters = i;
}

// poisson::FFT with the caller's face coefficient (poisson.cpp:391-411): each
// iteration forms the residual in its own kernel into `res` and runs the
// five-pass transform on it.  P, RHS, C and res are dense device arrays.
void do_solve_coef(tpx_ctx *c, const double *const cf[3], double *P, const double *rhs, double *res, int *iters) {
  const int i = solve_loop(c, 2, [&](int) {
    launch_residual_coef(c, cf, P, rhs, res);
    launch_poisson_iteration_on(c, nullptr, P, rhs, c->scal, res);
  });
  c->num_iteration = i;  // the reference's poisson object keeps the last solve's count
  if (iters) *iters = i;
}

// grid::Update_RV_WOQ (grid.cpp:646-692); U = RU_int/Rho_face from the stage state
void do_update_rv_woq(tpx_ctx *c, int rk) {
  ensure_passive_vector(c);
  const double *ri[3];
  ru_ptrs(c, c->ru_int, ri);
  launch_rv_woq(c, coef(c, rk), ri, c->rho[c->rho_int]);
  c->div_valid = false;  // the reference's Update_RV_WOQ overwrites divergence with div(V)
}

// grid::Compute_RHS_Pois_Q (grid.cpp:694-701); its mean is removed where the
// solve reads it (and on download)
void do_compute_rhs_pois_q(tpx_ctx *c, int rk) {
  ensure_passive_vector(c);
  launch_rhs_pois_q(c, coef(c, rk), c->stats_on && rk == 3);
}

// grid::Solve_Poisson_Q (grid.cpp:703-707): the FFT iteration with the
// constant density Rho_forV; the solver's iteration count is shared with the
// gas's solve, as the reference's single poisson object is
void do_solve_poisson_q(tpx_ctx *c, int *iters) {
  ensure_passive_vector(c);
  const int i = solve_loop(c, 2, [&](int) {
    launch_poisson_iteration_on(c, c->rho_forv_field, c->Qv, c->rhs_q, c->scal, nullptr, SC_SUM_RHS_Q);
  });
  c->num_iteration = i;
  if (iters) *iters = i;
}

// grid::Update_RV_WQ (grid.cpp:709-715)
void do_update_rv_wq(tpx_ctx *c, int rk) {
  ensure_passive_vector(c);
  launch_rv_wq(c, coef(c, rk));
}

void do_passive_vector(tpx_ctx *c, int rk) {
  do_update_rv_woq(c, rk);
  do_compute_rhs_pois_q(c, rk);
  do_solve_poisson_q(c, nullptr);
  do_update_rv_wq(c, rk);
}

void do_update_ru_wp(tpx_ctx *c, int rk, bool want_sums) {
  Coef k = coef(c, rk);
  double *on[3], *ow[3];
  ru_wptrs(c, c->ru_np1, on);
  ru_wptrs(c, c->ru_new, ow);
  launch_project(c, k, c->P, on, ow, rk == 3, want_sums);
}

void do_time_advance_rk4(tpx_ctx *c) {
  c->rho_int = c->rho_new;
  c->ru_int = c->ru_new;
  c->rho_new = -1;
  c->ru_new = -1;
  launch_scalar_copy(c, SC_P0_INT, SC_P0_NEW);
  if (c->ps[0]) std::swap(c->ps_int, c->ps_new);  // Passive_Scalar_int = Passive_Scalar_new (grid.cpp:897)
  if (c->Qv) std::swap(c->rv_int, c->rv_new);     // RV_int = RV_new (grid.cpp:894)
  Particles &pt = c->pt;
  std::swap(pt.st_int, pt.st_new);  // with int_is_base, new receives the spare buffer
  pt.int_is_base = false;
  std::swap(pt.perm_int, pt.perm_new);
  std::swap(pt.inv_int, pt.inv_new);
  std::swap(pt.start_int, pt.start_new);
}

void do_make_mean_u0(tpx_ctx *c, const double a[3], bool sums_ready) {
  int src = ru_np1_read(c);
  if (!sums_ready)
    for (int d = 0; d < 3; ++d) launch_reduce(c, c->ru[src][d], 0, SC_MEAN_RU0 + d);
  int dst = pick_free({c->ru_base, c->ru_int, c->ru_new, src}, 5);
  const double *in[3];
  double *out[3];
  ru_ptrs(c, src, in);
  ru_wptrs(c, dst, out);
  launch_shift_mean(c, in, out, a);
  c->ru_np1 = dst;
  c->ru_np1_is_base = false;
}

void do_time_advance(tpx_ctx *c) {
  c->rho_base = rho_np1_read(c);
  c->rho_np1_is_base = true;
  c->ru_base = ru_np1_read(c);
  c->ru_np1_is_base = true;
  launch_scalar_copy(c, SC_P0, SC_P0_NP1);
  if (c->ps[0]) launch_copy(c, c->ps[c->ps_base], c->ps[c->ps_np1], c->g.n);  // grid.cpp:598
  if (c->Qv)  // RV = RV_np1 (grid.cpp:600)
    for (int d = 0; d < 3; ++d) launch_copy(c, c->rv[c->rv_base][d], c->rv[c->rv_np1][d], c->g.n);
  c->T_cur += c->p.dt;
  c->num_timestep++;
  Particles &pt = c->pt;
  if (!pt.np1_is_base) std::swap(pt.st_base, pt.st_np1);
  pt.np1_is_base = true;
  if (pt.n > 0 && c->binned_this_step) {
    if (c->adj.step >= 0) adjoint_track_permutation(c);
    launch_particle_permute(c);
  }
  c->binned_this_step = false;
}

void do_params_update(tpx_ctx *c) {
  if (c->T_cur > c->p.threshold) {  // params.cpp:75-86
    c->Iteration = c->p.iteration2;
    c->A = c->p.a2;
    c->gv[0] = c->p.gx2; c->gv[1] = c->p.gy2; c->gv[2] = c->p.gz2;
    c->I0 = c->p.i02;
  }
}

// The fused substep: the reference's method sequence with Update_Rho moved
// after the particle advance (it reads nothing the particles write) and
// merged into the momentum kernel, and Compute_Div_U_new merged into
// Compute_RHS_Pois.  Results are identical to the method sequence.  The
// part before the Poisson solve:
void fused_substep_pre(tpx_ctx *c, int rk) {
  Coef k = coef(c, rk);
  launch_update_p0(c, k);
  Particles &pt = c->pt;
  bool parts = pt.n > 0;
  bool rebin = parts && (c->p.twc || heat_active(c));
  if (c->stats_on && rk == 3) {  // Update_Particle's concentration and gather at this stage
    const double *ri[3];
    ru_ptrs(c, c->ru_int, ri);
    stats_capture_particles(c, ri, c->rho[c->rho_int]);
  }
  if (parts) {
    const double *ri[3];
    ru_ptrs(c, c->ru_int, ri);
    if (c->adj.step >= 0) adjoint_tape_substep(c, rk, ri, c->rho[c->rho_int]);
    launch_particle_advance(c, k, ri, c->rho[c->rho_int], rebin);
    if (c->p.twc) launch_velocity_deposit(c);
  }
  if (c->p.solve_for_scalar) do_update_passive_scalar(c, rk);  // simulation.cpp:65
  if (c->p.solve_for_vector) do_passive_vector(c, rk);          // simulation.cpp:66-71
  do_update_ru_wop(c, rk, true);
  if (rebin) {
    launch_sort_new(c);
    c->binned_this_step = true;
  }
  if (parts && heat_active(c)) {
    launch_heat_exchange(c, TPX_NEW, c->rho[c->rho_new], SC_P0_NEW);
    c->q_zero = false;
  } else if (!c->q_zero) {
    if (parts) launch_fill(c, pt.rhs_t, 0.0, pt.n);
    launch_fill(c, c->Q, 0.0, c->g.n);
    launch_fill(c, c->scal + SC_SUM_Q, 0.0, 1);
    c->q_zero = true;
  }
  const double *rn[3];
  ru_ptrs(c, c->ru_new, rn);
  if (c->stats_on && rk == 3) stats_capture_velocity(c, rn, c->rho[c->rho_new]);  // Compute_RHS_Pois's U
  launch_div_rhs(c, k, 0, rn, c->rho[c->rho_new], c->q_zero ? nullptr : c->Q, nullptr, c->rhs);
  c->div_valid = false;
}

void fused_substep_post(tpx_ctx *c, int rk) {
  do_update_ru_wp(c, rk, rk == 3);
  do_time_advance_rk4(c);
}

void fused_substep(tpx_ctx *c, int rk) {
  fused_substep_pre(c, rk);
  do_solve_poisson(c, rk, nullptr);
  fused_substep_post(c, rk);
}

// the end of a step after its four substeps (simulation.cpp:80-82)
void step_end(tpx_ctx *c) {
  const double a[3] = {c->p.u0 * c->p.rho0, c->p.v0 * c->p.rho0, c->p.w0 * c->p.rho0};
  do_make_mean_u0(c, a, true);
  do_time_advance(c);
}

// frees every array of a recorded adjoint run
void adjoint_free(AdjointTape &t) {
  auto f = [](double *&p) { if (p) (void)hipFree(p); p = nullptr; };
  f(t.gas); f(t.ckpt); f(t.y); f(t.seg); f(t.lb); f(t.lnext); f(t.S); f(t.stage); f(t.lam); f(t.dtau); f(t.djdc);
  f(t.start); f(t.final_state); f(t.work); f(t.cap); f(t.lup); f(t.iters_seen); f(t.sub); f(t.tau0);
  f(t.qt); f(t.qs); f(t.p0t); f(t.eps0); f(t.deps); f(t.lh); f(t.hpart); f(t.scal_save); f(t.gck);
  t.heat = false;
  t.segmented = false;
  t.gas_base = 0;
  t.seg_loaded = -1;
  t.nq = 6;
  t.swept = false;
  auto fi = [](int32_t *&p) { if (p) (void)hipFree(p); p = nullptr; };
  fi(t.cum); fi(t.cum_tmp);
  f(t.lug); f(t.gt); f(t.segstage);
  t.segstage_tried = false;
  fi(t.bkey); fi(t.bperm); fi(t.binv); fi(t.bstart);
  t.store = false;
  for (auto &p : t.pcap) f(p);
  t.pcap.clear();
  t.ga = GasAdjoint{};
  t.ready = false;
}

void release(tpx_ctx *c) {
  if (!c) return;
  auto f = [](void *p) { if (p) (void)hipFree(p); };
  for (auto &b : c->ru) for (auto &p : b) f(p);
  for (auto &p : c->rho) f(p);
  f(c->P); f(c->rhs); f(c->Q); f(c->C); f(c->divergence); f(c->dep_tmp); f(c->coef_stage);
  for (auto &set : c->rv) for (auto &p : set) f(p);
  f(c->Qv); f(c->rhs_q); f(c->rho_forv_field); f(c->rv_filt); f(c->track);
  for (auto &p : c->S) f(p);
  for (auto &p : c->spec) f(p);
  for (auto &p : c->lam) f(p);
  for (auto &p : c->tw) f(p);
  f(c->scal); f(c->partials); f(c->scan_tmp); f(c->st_part);
  for (auto &p : c->ps) f(p);
  adjoint_free(c->adj);
  if (c->h_scal) (void)hipHostFree(c->h_scal);
  Particles &pt = c->pt;
  for (auto &q : pt.q) for (auto &p : q) f(p);
  f(pt.rhs_t); f(pt.tau); f(pt.eps); f(pt.flag); f(pt.contrib); f(pt.key);
  f(pt.perm_int); f(pt.perm_new); f(pt.start_int); f(pt.start_new); f(pt.cursor);
  f(pt.inv_int); f(pt.inv_new);
  for (auto &p : pt.spare) f(p);
  f(pt.itmp);
  for (auto &e : c->tile.pending) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
  for (auto &e : c->tm.pool) (void)hipEventDestroy(e);
  for (auto &r : c->tm.recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

std::string validate(const tpx_params *p) {
  if (!is_pow2(p->nx) || !is_pow2(p->ny) || !is_pow2(p->nz))
    return "grid sizes must be powers of two (FFT Poisson solve)";
  if (p->nx < 8 || p->ny < 8 || p->nz < 8 || p->nx > 512 || p->ny > 512 || p->nz > 512)
    return "grid sizes must lie in [8, 512]";
  if ((long)p->nx * p->ny * p->nz >= (1L << 31)) return "grid too large";
  if (!(p->lx > 0 && p->ly > 0 && p->lz > 0)) return "box lengths must be positive";
  if (!(p->dt > 0)) return "dt must be positive";
  if (p->bs < 0) return "bs must be >= 0";
  if (!(p->mu0 > 0)) return "Mu0 must be positive (it divides Tp)";
  return "";
}

}  // namespace

namespace tpx {

void timing_begin(tpx_ctx *c, int cls) {
  if (!c->tm.on || g_rec) return;
  Timing &t = c->tm;
  hipEvent_t a, b;
  if (t.pool.size() >= 2) {
    a = t.pool.back(); t.pool.pop_back();
    b = t.pool.back(); t.pool.pop_back();
  } else {
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
  }
  hip_check(hipEventRecord(a, c->stream), "hipEventRecord");
  t.recs.push_back({cls, a, b});
}

void timing_end(tpx_ctx *c, int cls) {
  if (!c->tm.on || g_rec) return;
  Timing &t = c->tm;
  for (auto it = t.recs.rbegin(); it != t.recs.rend(); ++it)
    if (it->cls == cls) {
      hip_check(hipEventRecord(it->b, c->stream), "hipEventRecord");
      return;
    }
}

}  // namespace tpx

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *tpx_create_error(void) { return g_create_error.c_str(); }

const char *tpx_last_error(const tpx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tpx_create(const tpx_params *p, int64_t np_capacity, tpx_ctx **out) {
  if (!p || !out) { g_create_error = "null argument"; return TPX_E_ARG; }
  *out = nullptr;
  std::string why = validate(p);
  if (!why.empty()) { g_create_error = why; return TPX_E_ARG; }
  if (np_capacity < 0 || np_capacity > (int64_t)INT32_MAX) {
    g_create_error = "np_capacity out of range";
    return TPX_E_ARG;
  }
  tpx_ctx *c = new tpx_ctx();
  try {
    c->p = *p;
    hip_check(hipGetDevice(&c->device), "hipGetDevice");
    c->static_scratch = static_scratch(c->device);
    int lds = 0;
    hip_check(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device),
              "hipDeviceGetAttribute");
    c->max_lds = (size_t)lds;
    // TPX_FFT_PASSES=5 selects the five-pass Poisson iteration everywhere
    // (the three-pass form is the default where the axis lengths allow it)
    if (const char *e = getenv("TPX_FFT_PASSES")) c->fft_passes = atoi(e) == 5 ? 5 : 3;
    // TPX_XPACK=0 keeps the full-length complex x transform at x lengths 64, 128 and 256
    if (const char *e = getenv("TPX_XPACK")) c->x_packed = atoi(e) != 0;
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    Geo &g = c->g;
    g.nx = p->nx; g.ny = p->ny; g.nz = p->nz;
    g.nxy = (long)g.nx * g.ny;
    g.n = g.nxy * g.nz;
    // gridsize.cpp:23-25, tensor0.cpp:12-32
    g.dx = p->lx / p->nx; g.dy = p->ly / p->ny; g.dz = p->lz / p->nz;
    g.idx = 1. / g.dx; g.idy = 1. / g.dy; g.idz = 1. / g.dz;
    g.idx2 = 1. / (g.dx * g.dx); g.idy2 = 1. / (g.dy * g.dy); g.idz2 = 1. / (g.dz * g.dz);
    c->Vcell = g.dx * g.dy * g.dz;
    // params.cpp:61-72
    c->R = p->cp - p->cv;
    c->P0par = p->rho0 * p->t0 * c->R;
    c->Tp = p->rhop * p->dp * p->dp / (18 * p->mu0);
    c->mp = p->rhop * TPX_PI * p->dp * p->dp * p->dp / 6.;
    c->Iteration = p->iteration1;
    c->A = p->a1;
    c->gv[0] = p->gx1; c->gv[1] = p->gy1; c->gv[2] = p->gz1;
    c->I0 = p->i01;
    c->pre[0] = 0.5; c->pre[1] = 0.5; c->pre[2] = 1.; c->pre[3] = 1. / 6.;
    c->post[0] = 1. / 6.; c->post[1] = 1. / 3.; c->post[2] = 1. / 3.; c->post[3] = 1. / 6.;
    auto dalloc = [&](double **ptr, size_t n) {
      if (hipMalloc(ptr, sizeof(double) * (n ? n : 1)) != hipSuccess) throw std::bad_alloc();
    };
    auto ialloc = [&](int32_t **ptr, size_t n) {
      if (hipMalloc(ptr, sizeof(int32_t) * (n ? n : 1)) != hipSuccess) throw std::bad_alloc();
    };
    size_t n = (size_t)g.n;
    for (auto &b : c->ru) for (auto &f : b) dalloc(&f, n);
    for (auto &f : c->rho) dalloc(&f, n);
    dalloc(&c->P, n); dalloc(&c->rhs, n); dalloc(&c->Q, n); dalloc(&c->C, n); dalloc(&c->divergence, n);
    if (np_capacity > 0) dalloc(&c->dep_tmp, 3 * n);
    for (auto &f : c->S) dalloc(&f, n);
    dalloc(&c->scal, SC_COUNT);
    // one partial per 256-cell block or per x-row, whichever is more
    c->n_partials = std::max((long)grid1d(g.n, 256), (long)g.ny * g.nz) + 1024;
    dalloc(&c->partials, 3 * (size_t)c->n_partials);
    ialloc(&c->scan_tmp, (size_t)(g.n / 2048 + 1024));
    if (hipHostMalloc((void **)&c->h_scal, sizeof(double) * SC_COUNT, hipHostMallocDefault) != hipSuccess)
      throw std::bad_alloc();
    memset(c->h_scal, 0, sizeof(double) * SC_COUNT);
    hip_check(hipMemset(c->scal, 0, sizeof(double) * SC_COUNT), "hipMemset");
    for (auto &f : c->S) hip_check(hipMemset(f, 0, sizeof(double) * n), "hipMemset");
    hip_check(hipMemset(c->Q, 0, sizeof(double) * n), "hipMemset");
    hip_check(hipMemset(c->P, 0, sizeof(double) * n), "hipMemset");
    int rc = fft_setup(c);
    if (rc != TPX_OK) throw std::runtime_error("FFT setup failed");
    // particles
    Particles &pt = c->pt;
    pt.cap = np_capacity;
    size_t cap = (size_t)np_capacity;
    for (auto &q : pt.q) for (auto &a : q) dalloc(&a, cap);
    dalloc(&pt.rhs_t, cap); dalloc(&pt.tau, cap); dalloc(&pt.eps, cap);
    for (auto &a : pt.spare) dalloc(&a, cap);
    ialloc(&pt.flag, cap); ialloc(&pt.itmp, cap);
    dalloc(&pt.contrib, 6 * cap);
    ialloc(&pt.key, cap);
    ialloc(&pt.perm_int, cap); ialloc(&pt.perm_new, cap);
    ialloc(&pt.inv_int, cap); ialloc(&pt.inv_new, cap);
    ialloc(&pt.start_int, n + 1); ialloc(&pt.start_new, n + 1); ialloc(&pt.cursor, n + 1);
    hip_check(hipMemset(pt.start_int, 0, sizeof(int32_t) * (n + 1)), "hipMemset");
    hip_check(hipMemset(pt.rhs_t, 0, sizeof(double) * (cap ? cap : 1)), "hipMemset");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    // buffer roles: int aliases base until the first substep writes new
    c->ru_base = 0; c->ru_int = 0; c->ru_new = -1; c->ru_np1 = -1; c->ru_np1_is_base = true;
    c->rho_base = 0; c->rho_int = 0; c->rho_new = -1; c->rho_np1 = -1; c->rho_np1_is_base = true;
    c->q_zero = true;
  } catch (const std::bad_alloc &) {
    g_create_error = "device allocation failed";
    release(c);
    return TPX_E_OOM;
  } catch (const std::exception &e) {
    g_create_error = e.what();
    release(c);
    return TPX_E_HIP;
  }
  *out = c;
  return TPX_OK;
}

int tpx_destroy(tpx_ctx *ctx) {
  if (!ctx) return TPX_E_ARG;
  (void)hipStreamSynchronize(ctx->stream);
  release(ctx);
  return TPX_OK;
}

int tpx_upload_field(tpx_ctx *ctx, int field, int comp, const double *host, int bs) {
  return guarded(ctx, [&]() {
    if (!host || bs < 0) throw std::invalid_argument("bad host array or bs");
    if (is_vector(field) ? (comp < 0 || comp > 2) : comp != 0) throw std::invalid_argument("bad component");
    if (!(field == TPX_RU || field == TPX_RHO || field == TPX_P || field == TPX_RHS_PART_TEMP ||
          field == TPX_RHS_RU || field == TPX_PASSIVE_SCALAR || field == TPX_RV || field == TPX_Q))
      throw std::invalid_argument("only RU, Rho, P, RHS_Part_Temp, RHS_RU, Passive_Scalar, RV and Q can be uploaded");
    const Geo &g = ctx->g;
    // uploading the base state resets the aliasing of int/np1 to the base
    if (field == TPX_RU && ctx->ru_int != ctx->ru_base) throw std::logic_error("upload RU before the first substep only");
    if (field == TPX_RHO && ctx->rho_int != ctx->rho_base) throw std::logic_error("upload Rho before the first substep only");
    double *dst = field_buffer(ctx, field, comp);
    std::vector<double> tmp((size_t)g.n);
    int GX = g.nx + 2 * bs, GY = g.ny + 2 * bs;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i)
          tmp[((size_t)k * g.ny + j) * g.nx + i] = host[((size_t)(k + bs) * GY + (j + bs)) * GX + (i + bs)];
    hip_check(hipMemcpyAsync(dst, tmp.data(), sizeof(double) * g.n, hipMemcpyHostToDevice, ctx->stream), "upload");
    if (field == TPX_RHS_PART_TEMP) {
      launch_reduce(ctx, ctx->Q, 0, SC_SUM_Q);
      ctx->q_zero = false;
    }
    sync(ctx);
    return TPX_OK;
  });
}

int tpx_download_field(tpx_ctx *ctx, int field, int comp, double *host, int bs) {
  return guarded(ctx, [&]() {
    if (!host || bs < 0) throw std::invalid_argument("bad host array or bs");
    if (is_vector(field) ? (comp < 0 || comp > 2) : comp != 0) throw std::invalid_argument("bad component");
    const double *src = field_buffer(ctx, field, comp);
    if (!src) throw std::logic_error("field not available in the current state");
    const Geo &g = ctx->g;
    std::vector<double> tmp((size_t)g.n);
    hip_check(hipMemcpyAsync(tmp.data(), src, sizeof(double) * g.n, hipMemcpyDeviceToHost, ctx->stream), "download");
    sync(ctx);
    if (field == TPX_RHS_POIS || field == TPX_RHS_POIS_Q) {  // make_mean_zero applied where the solve reads it
      double mean = read_scalar(ctx, field == TPX_RHS_POIS ? SC_SUM_RHS : SC_SUM_RHS_Q) / (double)(g.n);
      for (auto &v : tmp) v -= mean;
    }
    int GX = g.nx + 2 * bs, GY = g.ny + 2 * bs, GZ = g.nz + 2 * bs;
    for (int k = 0; k < GZ; ++k)
      for (int j = 0; j < GY; ++j)
        for (int i = 0; i < GX; ++i) {
          int si = ((i - bs) % g.nx + g.nx) % g.nx, sj = ((j - bs) % g.ny + g.ny) % g.ny,
              sk = ((k - bs) % g.nz + g.nz) % g.nz;
          host[((size_t)k * GY + j) * GX + i] = tmp[((size_t)sk * g.ny + sj) * g.nx + si];
        }
    return TPX_OK;
  });
}

int tpx_upload_particles(tpx_ctx *ctx, int64_t n, const double *x, const double *y, const double *z,
                         const double *u, const double *v, const double *w, const double *T,
                         const int32_t *flag, const double *tau, const double *eps) {
  return guarded(ctx, [&]() {
    Particles &pt = ctx->pt;
    if (n < 0 || n > pt.cap) throw std::invalid_argument("particle count exceeds the capacity");
    if (n > 0 && (!x || !y || !z || !u || !v || !w || !T || !flag)) throw std::invalid_argument("null particle array");
    const double *src[7] = {x, y, z, u, v, w, T};
    // the binning and the deposits assume every particle inside the box, as
    // the periodic wrap keeps it after every substep (particle.cpp:58-104)
    const double Lb[3] = {ctx->p.lx, ctx->p.ly, ctx->p.lz};
    for (int d = 0; d < 3; ++d)
      for (int64_t e = 0; e < n; ++e)
        if (!(src[d][e] >= 0.0 && src[d][e] <= Lb[d]))
          throw std::invalid_argument("particle position outside the box [0, L]");
    pt.st_base = 0; pt.st_int = 1; pt.st_new = 2; pt.st_np1 = 3;
    pt.np1_is_base = true;
    pt.int_is_base = false;
    for (int q = 0; q < 7; ++q)
      if (n) hip_check(hipMemcpyAsync(pt.q[q][pt.st_base], src[q], sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream), "upload");
    if (n) hip_check(hipMemcpyAsync(pt.flag, flag, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream), "upload");
    pt.has_tau = tau != nullptr;
    pt.has_eps = eps != nullptr;
    if (tau && n) hip_check(hipMemcpyAsync(pt.tau, tau, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream), "upload");
    if (eps && n) hip_check(hipMemcpyAsync(pt.eps, eps, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream), "upload");
    pt.n = n;
    pt.new_counted = false;
    sync(ctx);
    return TPX_OK;
  });
}

int tpx_num_particles(tpx_ctx *ctx, int64_t *n) {
  return guarded(ctx, [&]() {
    if (!n) throw std::invalid_argument("null");
    *n = ctx->pt.n;
    return TPX_OK;
  });
}

int tpx_download_particles(tpx_ctx *ctx, int stage, double *x, double *y, double *z, double *u, double *v,
                           double *w, double *T, int32_t *flag) {
  return guarded(ctx, [&]() {
    Particles &pt = ctx->pt;
    int s;
    switch (stage) {
    case TPX_BASE: s = pt.st_base; break;
    case TPX_INT: s = pt.int_buf(); break;
    case TPX_NEW: s = pt.st_new; break;
    case TPX_NP1: s = pt.np1_is_base ? pt.st_base : pt.st_np1; break;
    default: throw std::invalid_argument("bad stage");
    }
    double *dst[7] = {x, y, z, u, v, w, T};
    for (int q = 0; q < 7; ++q)
      if (dst[q] && pt.n)
        hip_check(hipMemcpyAsync(dst[q], pt.q[q][s], sizeof(double) * pt.n, hipMemcpyDeviceToHost, ctx->stream), "download");
    if (flag && pt.n) hip_check(hipMemcpyAsync(flag, pt.flag, sizeof(int32_t) * pt.n, hipMemcpyDeviceToHost, ctx->stream), "download");
    sync(ctx);
    return TPX_OK;
  });
}

int tpx_download_particle_rhs_t(tpx_ctx *ctx, double *rhs_t) {
  return guarded(ctx, [&]() {
    if (!rhs_t) throw std::invalid_argument("null");
    if (ctx->pt.n)
      hip_check(hipMemcpyAsync(rhs_t, ctx->pt.rhs_t, sizeof(double) * ctx->pt.n, hipMemcpyDeviceToHost, ctx->stream), "download");
    sync(ctx);
    return TPX_OK;
  });
}

int tpx_set_restart_numbers(tpx_ctx *ctx, double P0, double T_cur, int num_timestep) {
  return guarded(ctx, [&]() {
    write_scalar(ctx, SC_P0, P0);
    ctx->T_cur = T_cur;
    ctx->num_timestep = num_timestep;
    return TPX_OK;
  });
}

int tpx_get_scalars(tpx_ctx *ctx, double *o) {
  return guarded(ctx, [&]() {
    if (!o) throw std::invalid_argument("null");
    hip_check(hipMemcpyAsync(ctx->h_scal, ctx->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, ctx->stream), "download");
    sync(ctx);
    const double *h = ctx->h_scal;
    o[TPX_S_P0] = h[SC_P0]; o[TPX_S_P0_INT] = h[SC_P0_INT]; o[TPX_S_P0_NEW] = h[SC_P0_NEW];
    o[TPX_S_P0_NP1] = h[SC_P0_NP1]; o[TPX_S_T_CUR] = ctx->T_cur; o[TPX_S_NUM_TIMESTEP] = ctx->num_timestep;
    o[TPX_S_NUM_ITERATION] = ctx->num_iteration; o[TPX_S_MEAN_ENERGY] = h[SC_MET]; o[TPX_S_A] = ctx->A;
    o[TPX_S_I0] = ctx->I0; o[TPX_S_ITERATION] = ctx->Iteration;
    o[TPX_S_GX] = ctx->gv[0]; o[TPX_S_GY] = ctx->gv[1]; o[TPX_S_GZ] = ctx->gv[2];
    o[TPX_S_DP0_DT] = h[SC_DP0_DT];
    o[TPX_S_TOTAL_ITERATIONS] = (double)ctx->total_iterations;
    return TPX_OK;
  });
}

// G14: prep tail of grid::Initialize (grid.cpp:460-486)
int tpx_initialize(tpx_ctx *ctx) {
  return guarded(ctx, [&]() {
    tpx_ctx *c = ctx;
    c->ru_int = c->ru_base; c->ru_new = -1; c->ru_np1 = -1; c->ru_np1_is_base = true;
    c->rho_int = c->rho_base; c->rho_new = -1; c->rho_np1 = -1; c->rho_np1_is_base = true;
    launch_scalar_copy(c, SC_P0_INT, SC_P0);
    launch_scalar_copy(c, SC_P0_NP1, SC_P0);
    if (c->p.solve_for_scalar) ensure_passive_scalar(c);
    if (c->ps[0]) {  // grid.cpp:465-466 (the face values are formed from int where used)
      launch_copy(c, c->ps[c->ps_int], c->ps[c->ps_base], c->g.n);
      launch_copy(c, c->ps[c->ps_np1], c->ps[c->ps_base], c->g.n);
    }
    if (c->p.solve_for_scalar) launch_passive_scalar_source_sum(c);
    if (c->p.solve_for_vector) ensure_passive_vector(c);
    if (c->Qv) {  // RV_int = RV, RV_np1 = RV (grid.cpp:463-464)
      for (int d = 0; d < 3; ++d) {
        launch_copy(c, c->rv[c->rv_int][d], c->rv[c->rv_base][d], c->g.n);
        launch_copy(c, c->rv[c->rv_np1][d], c->rv[c->rv_base][d], c->g.n);
      }
    }
    if (c->p.solve_for_vector && c->p.s2_type == 0) launch_rv_source_sum(c);
    Particles &pt = c->pt;
    pt.np1_is_base = true;
    if (pt.n > 0) {
      // bin the particles by cell and store them in that order (base and int)
      launch_sort_stage(c, TPX_BASE, true);
      launch_particle_permute(c);
    }
    if (pt.n > 0 && heat_active(c)) {
      launch_heat_exchange(c, TPX_INT, c->rho[c->rho_int], SC_P0_INT);
      c->q_zero = false;
    } else {
      if (pt.n > 0) launch_fill(c, pt.rhs_t, 0.0, pt.n);
      launch_fill(c, c->Q, 0.0, c->g.n);
      launch_fill(c, c->scal + SC_SUM_Q, 0.0, 1);
      c->q_zero = true;
    }
    sync(c);
    return TPX_OK;
  });
}

int tpx_params_update(tpx_ctx *ctx) {
  return guarded(ctx, [&]() { do_params_update(ctx); return TPX_OK; });
}

#define TPX_RK_CHECK(rk) \
  if ((rk) < 0 || (rk) > 3) throw std::invalid_argument("RK4_count must be 0..3")

int tpx_update_rho(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() { TPX_RK_CHECK(rk); do_update_rho(ctx, rk); return TPX_OK; });
}

int tpx_update_p0(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() { TPX_RK_CHECK(rk); launch_update_p0(ctx, coef(ctx, rk)); return TPX_OK; });
}

int tpx_update_particle(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (ctx->rho_new < 0) throw std::logic_error("Update_Particle needs Rho_new (call Update_Rho first)");
    do_update_particle(ctx, rk);
    return TPX_OK;
  });
}

int tpx_update_passive_scalar(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (!ctx->p.solve_for_scalar) throw std::logic_error("Update_Passive_Scalar with solve_for_scalar = 0");
    do_update_passive_scalar(ctx, rk);
    return TPX_OK;
  });
}

int tpx_update_rv_woq(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (!ctx->p.solve_for_vector) throw std::logic_error("Update_RV_WOQ with solve_for_vector = 0");
    do_update_rv_woq(ctx, rk);
    return TPX_OK;
  });
}

int tpx_compute_rhs_pois_q(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (!ctx->p.solve_for_vector) throw std::logic_error("Compute_RHS_Pois_Q with solve_for_vector = 0");
    do_compute_rhs_pois_q(ctx, rk);
    return TPX_OK;
  });
}

int tpx_solve_poisson_q(tpx_ctx *ctx, int rk, int *iterations) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (!ctx->p.solve_for_vector) throw std::logic_error("Solve_Poisson_Q with solve_for_vector = 0");
    do_solve_poisson_q(ctx, iterations);
    return TPX_OK;
  });
}

int tpx_update_rv_wq(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (!ctx->p.solve_for_vector) throw std::logic_error("Update_RV_WQ with solve_for_vector = 0");
    do_update_rv_wq(ctx, rk);
    return TPX_OK;
  });
}

int tpx_update_ru_wop(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    do_update_ru_wop(ctx, rk, false);
    ctx->div_valid = false;  // the reference's Update_RU_WOP overwrites divergence with div(U)
    return TPX_OK;
  });
}

int tpx_compute_div_u_new(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (ctx->rho_new < 0) throw std::logic_error("Compute_Div_U_new needs Rho_new");
    const double *rn[3] = {nullptr, nullptr, nullptr};
    launch_div_rhs(ctx, coef(ctx, rk), 1, rn, ctx->rho[ctx->rho_new], ctx->q_zero ? nullptr : ctx->Q,
                   ctx->divergence, nullptr);
    ctx->div_valid = true;
    return TPX_OK;
  });
}

int tpx_compute_rhs_pois(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (ctx->rho_new < 0 || ctx->ru_new < 0) throw std::logic_error("Compute_RHS_Pois needs RU_new and Rho_new");
    const double *rn[3];
    ru_ptrs(ctx, ctx->ru_new, rn);
    if (ctx->stats_on && rk == 3) stats_capture_velocity(ctx, rn, ctx->rho[ctx->rho_new]);
    launch_div_rhs(ctx, coef(ctx, rk), 2, rn, ctx->rho[ctx->rho_new], nullptr, ctx->divergence, ctx->rhs);
    return TPX_OK;
  });
}

int tpx_solve_poisson(tpx_ctx *ctx, int rk, int *iterations) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (ctx->rho_new < 0) throw std::logic_error("Solve_Poisson needs Rho_new");
    do_solve_poisson(ctx, rk, iterations);
    return TPX_OK;
  });
}

int tpx_poisson_solve(tpx_ctx *ctx, const double *cx, const double *cy, const double *cz, double *P,
                      const double *rhs, int bs, int *iterations) {
  return guarded(ctx, [&]() {
    if (!cx || !cy || !cz || !P || !rhs || bs < 0) throw std::invalid_argument("bad host array or bs");
    const tpx::Geo &g = ctx->g;
    const size_t n = (size_t)g.n;
    if (!ctx->coef_stage && hipMalloc((void **)&ctx->coef_stage, sizeof(double) * 6 * n) != hipSuccess)
      throw std::bad_alloc();
    double *d = ctx->coef_stage;
    const int GX = g.nx + 2 * bs, GY = g.ny + 2 * bs;
    std::vector<double> tmp(5 * n);
    const double *src[5] = {cx, cy, cz, P, rhs};
    for (int f = 0; f < 5; ++f)
      for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
          for (int i = 0; i < g.nx; ++i)
            tmp[f * n + ((size_t)k * g.ny + j) * g.nx + i] = src[f][((size_t)(k + bs) * GY + (j + bs)) * GX + (i + bs)];
    hip_check(hipMemcpyAsync(d, tmp.data(), sizeof(double) * 5 * n, hipMemcpyHostToDevice, ctx->stream), "upload");
    const double *cf[3] = {d, d + n, d + 2 * n};
    do_solve_coef(ctx, cf, d + 3 * n, d + 4 * n, d + 5 * n, iterations);
    hip_check(hipMemcpyAsync(tmp.data(), d + 3 * n, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream),
              "download");
    sync(ctx);
    const int GZ = g.nz + 2 * bs;
    for (int k = 0; k < GZ; ++k)
      for (int j = 0; j < GY; ++j)
        for (int i = 0; i < GX; ++i) {
          const int si = ((i - bs) % g.nx + g.nx) % g.nx, sj = ((j - bs) % g.ny + g.ny) % g.ny,
                    sk = ((k - bs) % g.nz + g.nz) % g.nz;
          P[((size_t)k * GY + j) * GX + i] = tmp[((size_t)sk * g.ny + sj) * g.nx + si];
        }
    return TPX_OK;
  });
}

int tpx_poisson_solve_device(tpx_ctx *ctx, const double *cx, const double *cy, const double *cz, double *P,
                             const double *rhs, int *iterations) {
  return guarded(ctx, [&]() {
    if (!cx || !cy || !cz || !P || !rhs) throw std::invalid_argument("null device array");
    if (ctx->g.n > kStaticScratchDoubles)
      throw std::invalid_argument("tpx_poisson_solve_device: the grid has " + std::to_string(ctx->g.n) +
                                  " cells, more than its static scratch holds (" +
                                  std::to_string(kStaticScratchDoubles) + " cells, 128^3)");
    if (!ctx->static_scratch) throw std::logic_error("static scratch not set up");
    const double *cf[3] = {cx, cy, cz};
    do_solve_coef(ctx, cf, P, rhs, ctx->static_scratch, iterations);
    return TPX_OK;
  });
}

int tpx_update_ru_wp(tpx_ctx *ctx, int rk) {
  return guarded(ctx, [&]() {
    TPX_RK_CHECK(rk);
    if (ctx->ru_new < 0 || ctx->ru_np1_is_base) throw std::logic_error("Update_RU_WP needs RU_new/RU_np1");
    do_update_ru_wp(ctx, rk, false);
    return TPX_OK;
  });
}

int tpx_time_advance_rk4(tpx_ctx *ctx) {
  return guarded(ctx, [&]() {
    if (ctx->ru_new < 0 || ctx->rho_new < 0) throw std::logic_error("TimeAdvance_RK4 before a substep");
    do_time_advance_rk4(ctx);
    return TPX_OK;
  });
}

int tpx_make_mean_u0(tpx_ctx *ctx, const double a[3]) {
  return guarded(ctx, [&]() {
    if (!a) throw std::invalid_argument("null");
    do_make_mean_u0(ctx, a, false);
    return TPX_OK;
  });
}

int tpx_time_advance(tpx_ctx *ctx) {
  return guarded(ctx, [&]() { do_time_advance(ctx); sync(ctx); return TPX_OK; });
}

int tpx_step(tpx_ctx *ctx, int nsteps) {
  return guarded(ctx, [&]() {
    if (nsteps < 0) throw std::invalid_argument("nsteps < 0");
    for (int s = 0; s < nsteps; ++s) {
      do_params_update(ctx);
      for (int rk = 0; rk < 4; ++rk) fused_substep(ctx, rk);
      step_end(ctx);
      if (ctx->tile.active) tile_tune_collect(ctx);
    }
    sync(ctx);
    return TPX_OK;
  });
}

int tpx_ensemble_step(tpx_ctx *const *ctxs, int n, int nsteps, int concurrency) {
  if (!ctxs || n < 0 || nsteps < 0) return TPX_E_ARG;
  std::vector<tpx_ctx *> seen(ctxs, ctxs + n);
  std::sort(seen.begin(), seen.end());
  for (int i = 0; i < n; ++i)
    if (!seen[i] || (i > 0 && seen[i] == seen[i - 1])) return TPX_E_ARG;  // null or repeated
  const int workers = std::max(1, std::min(n, concurrency > 0 ? concurrency : 16));
  std::atomic<int> next{0};
  std::atomic<int> first_err{TPX_OK};
  auto work = [&]() {
    for (int idx = next.fetch_add(1); idx < n; idx = next.fetch_add(1)) {
      tpx_ctx *c = ctxs[idx];
      int rc = hipSetDevice(c->device) == hipSuccess ? tpx_step(c, nsteps) : TPX_E_HIP;
      if (rc != TPX_OK) {
        int expected = TPX_OK;
        first_err.compare_exchange_strong(expected, rc);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < workers; ++w) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  return first_err.load();
}

}  // extern "C"

// ---- batches: one launch per kernel for many boxes ------------------------
// The host runs each box's step logic with its launches recorded (buffer
// roles and flags advance per box exactly as in tpx_step), then starts every
// recorded kernel once for all boxes on the batch's stream, with the boxes'
// arguments packed into a table in device memory.  The Poisson loop test is
// per box on the device; the host queues iterations for all boxes until every
// box has stopped.
struct tpx_batch {
  std::vector<tpx_ctx *> ctx;
  std::vector<Recorder> rec;
  hipStream_t stream = nullptr;
  int device = 0;
  static constexpr int kRing = 4;  // argument tables in flight
  char *h_tab[kRing] = {};
  char *d_tab[kRing] = {};
  size_t cap[kRing] = {};
  hipEvent_t ev[kRing] = {};
  int slot = 0;
  double *d_state = nullptr, *h_state = nullptr;  // per box: iterations run, stopped
  // one Poisson iteration of every box as a captured graph, replayed for each
  // iteration of each solve; its argument table lives at a fixed address and
  // is refreshed (stream-ordered) at the start of each solve
  char *h_iter = nullptr, *d_iter = nullptr;
  size_t iter_cap = 0;
  hipGraphExec_t iter_exec = nullptr;
  std::vector<char> iter_sig;  // the launches the graph was captured from
  // the segments between the solves (a substep's kernels before and after
  // its solve, the step's end) as captured graphs, one per distinct launch
  // sequence; each keeps its argument table at a fixed device address,
  // refreshed (stream-ordered) before each replay
  struct Seg {
    std::vector<char> sig;
    char *h = nullptr, *d = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t copied = nullptr;  // the last table copy has read h
  };
  std::vector<Seg> segs;
  bool graphs = true;  // TPX_BATCH_GRAPH=0: launch the segments kernel by kernel
  std::string err;
};

namespace {

thread_local std::string g_batch_error;

void batch_release(tpx_batch *b) {
  if (!b) return;
  for (int i = 0; i < tpx_batch::kRing; ++i) {
    if (b->h_tab[i]) (void)hipHostFree(b->h_tab[i]);
    if (b->d_tab[i]) (void)hipFree(b->d_tab[i]);
    if (b->ev[i]) (void)hipEventDestroy(b->ev[i]);
  }
  if (b->d_state) (void)hipFree(b->d_state);
  if (b->h_state) (void)hipHostFree(b->h_state);
  if (b->iter_exec) (void)hipGraphExecDestroy(b->iter_exec);
  for (auto &e : b->segs) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.h) (void)hipHostFree(e.h);
    if (e.d) (void)hipFree(e.d);
    if (e.copied) (void)hipEventDestroy(e.copied);
  }
  if (b->h_iter) (void)hipHostFree(b->h_iter);
  if (b->d_iter) (void)hipFree(b->d_iter);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// A recorded segment: each kernel once for all boxes, with its place in the
// packed argument table.
struct BatchPlan {
  struct K {
    void (*launch)(dim3, dim3, size_t, hipStream_t, const char *, unsigned, unsigned);
    dim3 grid, block;
    size_t shmem, off, stride;
  };
  std::vector<K> k;
  size_t total = 0;
};

// record fn(ctx) for every box and lay out the table
template <class Fn>
void batch_record(tpx_batch *b, Fn &&fn, BatchPlan &plan) {
  const int nb = (int)b->ctx.size();
  for (int i = 0; i < nb; ++i) {
    b->rec[i].clear();
    b->rec[i].nbox = (unsigned)nb;
    g_rec = &b->rec[i];
    try {
      fn(b->ctx[i], i);
    } catch (...) {
      g_rec = nullptr;
      throw;
    }
    g_rec = nullptr;
  }
  const Recorder &r0 = b->rec[0];
  const size_t nk = r0.recs.size();
  plan.k.clear();
  plan.total = 0;
  for (size_t j = 0; j < nk; ++j) {
    const Recorder::Rec &a = r0.recs[j];
    BatchPlan::K k{a.launch, a.grid, a.block, a.shmem, 0, 0};
    for (int i = 1; i < nb; ++i) {
      if (b->rec[i].recs.size() != nk) throw std::logic_error("boxes of a batch launch different kernel sequences");
      const Recorder::Rec &o = b->rec[i].recs[j];
      if (o.launch != a.launch || o.block.x != a.block.x || o.block.y != a.block.y || o.block.z != a.block.z ||
          o.shmem != a.shmem || o.grid.z != a.grid.z)
        throw std::logic_error("boxes of a batch launch different kernel sequences");
      // particle kernels: the largest box's grid, the others' extra blocks exit
      k.grid.x = std::max(k.grid.x, o.grid.x);
      k.grid.y = std::max(k.grid.y, o.grid.y);
    }
    k.stride = (a.size + 15) / 16 * 16;
    k.off = plan.total;
    plan.total += (k.stride * nb + 255) / 256 * 256;
    plan.k.push_back(k);
  }
}

void batch_pack(const tpx_batch *b, const BatchPlan &plan, char *h) {
  const int nb = (int)b->ctx.size();
  for (size_t j = 0; j < plan.k.size(); ++j)
    for (int i = 0; i < nb; ++i) {
      const Recorder::Rec &o = b->rec[i].recs[j];
      memcpy(h + plan.k[j].off + (size_t)i * plan.k[j].stride, b->rec[i].bytes.data() + o.off, o.size);
    }
}

void batch_launch(tpx_batch *b, const BatchPlan &plan, const char *d) {
  const unsigned nb = (unsigned)b->ctx.size();
  for (const BatchPlan::K &k : plan.k) {
    k.launch(k.grid, k.block, k.shmem, b->stream, d + k.off, (unsigned)k.stride, nb);
    TPX_LAUNCH_CHECK();
  }
}

// record fn(ctx) for every box, then launch each recorded kernel once for all
// boxes, the arguments from one table copied in one transfer
template <class Fn>
void batch_run(tpx_batch *b, Fn &&fn) {
  BatchPlan plan;
  batch_record(b, fn, plan);
  if (plan.k.empty()) return;
  const int s = b->slot;
  b->slot = (s + 1) % tpx_batch::kRing;
  hip_check(hipEventSynchronize(b->ev[s]), "hipEventSynchronize");  // the slot's last copy has been read
  if (plan.total > b->cap[s]) {
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    if (b->h_tab[s]) hip_check(hipHostFree(b->h_tab[s]), "hipHostFree");
    if (b->d_tab[s]) hip_check(hipFree(b->d_tab[s]), "hipFree");
    b->h_tab[s] = nullptr;
    b->d_tab[s] = nullptr;
    b->cap[s] = 0;
    const size_t want = std::max<size_t>(plan.total * 2, 1 << 16);
    hip_check(hipHostMalloc((void **)&b->h_tab[s], want), "hipHostMalloc");
    if (hipMalloc((void **)&b->d_tab[s], want) != hipSuccess) throw std::bad_alloc();
    b->cap[s] = want;
  }
  batch_pack(b, plan, b->h_tab[s]);
  hip_check(hipMemcpyAsync(b->d_tab[s], b->h_tab[s], plan.total, hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync");
  hip_check(hipEventRecord(b->ev[s], b->stream), "hipEventRecord");
  batch_launch(b, plan, b->d_tab[s]);
}

// the launch sequence of a plan (what a graph captured from it depends on)
std::vector<char> plan_signature(const BatchPlan &plan) {
  std::vector<char> sig(sizeof(BatchPlan::K) * plan.k.size() + sizeof(size_t));
  for (size_t j = 0; j < plan.k.size(); ++j) {
    BatchPlan::K k = plan.k[j];
    memcpy(sig.data() + j * sizeof(k), &k, sizeof(k));
  }
  memcpy(sig.data() + sizeof(BatchPlan::K) * plan.k.size(), &plan.total, sizeof(size_t));
  return sig;
}

// batch_run as a graph replay: the segment's graph (captured the first time
// its launch sequence occurs) launched after a copy of this occurrence's
// argument table to the graph's fixed table
template <class Fn>
void batch_run_graph(tpx_batch *b, Fn &&fn) {
  if (!b->graphs) {
    batch_run(b, fn);
    return;
  }
  BatchPlan plan;
  batch_record(b, fn, plan);
  if (plan.k.empty()) return;
  std::vector<char> sig = plan_signature(plan);
  tpx_batch::Seg *e = nullptr;
  for (auto &x : b->segs)
    if (x.sig == sig) e = &x;
  if (!e) {
    // a new launch sequence: its table and graph, or nothing cached on failure
    tpx_batch::Seg n;
    n.sig = sig;
    hipGraph_t graph = nullptr;
    bool capturing = false;
    try {
      hip_check(hipHostMalloc((void **)&n.h, plan.total), "hipHostMalloc");
      if (hipMalloc((void **)&n.d, plan.total) != hipSuccess) throw std::bad_alloc();
      hip_check(hipEventCreateWithFlags(&n.copied, hipEventDisableTiming), "hipEventCreate");
      hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      capturing = true;
      batch_launch(b, plan, n.d);
      capturing = false;
      hip_check(hipStreamEndCapture(b->stream, &graph), "hipStreamEndCapture");
      const hipError_t err = hipGraphInstantiate(&n.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      graph = nullptr;
      hip_check(err, "hipGraphInstantiate");
    } catch (...) {
      if (capturing) (void)hipStreamEndCapture(b->stream, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      if (n.exec) (void)hipGraphExecDestroy(n.exec);
      if (n.h) (void)hipHostFree(n.h);
      if (n.d) (void)hipFree(n.d);
      if (n.copied) (void)hipEventDestroy(n.copied);
      throw;
    }
    b->segs.push_back(std::move(n));
    e = &b->segs.back();
  }
  hip_check(hipEventSynchronize(e->copied), "hipEventSynchronize");  // the last copy from e->h is done
  batch_pack(b, plan, e->h);
  hip_check(hipMemcpyAsync(e->d, e->h, plan.total, hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync");
  hip_check(hipEventRecord(e->copied, b->stream), "hipEventRecord");
  hip_check(hipGraphLaunch(e->exec, b->stream), "hipGraphLaunch");
}

// The Poisson iteration of every box, its table refreshed for this solve and
// the graph of its launches captured if the launches changed.
void batch_iteration_graph(tpx_batch *b) {
  BatchPlan plan;
  batch_record(b, [](tpx_ctx *c, int) { launch_poisson_iteration(c, c->rho[c->rho_new]); }, plan);
  if (plan.total > b->iter_cap) {
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    if (b->iter_exec) (void)hipGraphExecDestroy(b->iter_exec);
    b->iter_exec = nullptr;
    if (b->h_iter) hip_check(hipHostFree(b->h_iter), "hipHostFree");
    if (b->d_iter) hip_check(hipFree(b->d_iter), "hipFree");
    b->h_iter = nullptr;
    b->d_iter = nullptr;
    b->iter_cap = 0;
    const size_t want = std::max<size_t>(plan.total * 2, 1 << 16);
    hip_check(hipHostMalloc((void **)&b->h_iter, want), "hipHostMalloc");
    if (hipMalloc((void **)&b->d_iter, want) != hipSuccess) throw std::bad_alloc();
    b->iter_cap = want;
  }
  // the previous solve's copy has completed: batch_solve waits on the stream
  // before it returns
  batch_pack(b, plan, b->h_iter);
  hip_check(hipMemcpyAsync(b->d_iter, b->h_iter, plan.total, hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync");
  std::vector<char> sig(sizeof(BatchPlan::K) * plan.k.size());
  for (size_t j = 0; j < plan.k.size(); ++j) {
    BatchPlan::K k = plan.k[j];
    memcpy(sig.data() + j * sizeof(k), &k, sizeof(k));
  }
  if (b->iter_exec && sig == b->iter_sig) return;
  if (b->iter_exec) {
    (void)hipGraphExecDestroy(b->iter_exec);
    b->iter_exec = nullptr;
  }
  hipGraph_t graph = nullptr;
  hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    batch_launch(b, plan, b->d_iter);
  } catch (...) {
    (void)hipStreamEndCapture(b->stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  hip_check(hipStreamEndCapture(b->stream, &graph), "hipStreamEndCapture");
  const hipError_t e = hipGraphInstantiate(&b->iter_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  hip_check(e, "hipGraphInstantiate");
  b->iter_sig = sig;
}

void batch_solve(tpx_batch *b, int rk) {
  const int nb = (int)b->ctx.size();
  int itmax = 0, guess = 1;
  for (tpx_ctx *c : b->ctx) {
    itmax = std::max(itmax, c->Iteration);
    guess = std::max(guess, c->iter_guess[rk & 3]);
  }
  batch_run(b, [](tpx_ctx *c, int) { launch_solve_begin(c); });
  batch_iteration_graph(b);
  int queued = 0, batch = guess;
  while (queued < itmax) {
    const int n = std::min(batch, itmax - queued);
    for (int q = 0; q < n; ++q) hip_check(hipGraphLaunch(b->iter_exec, b->stream), "hipGraphLaunch");
    queued += n;
    batch_run(b, [&](tpx_ctx *c, int i) { launch_solve_state(c, b->d_state + 2 * i); });
    hip_check(hipMemcpyAsync(b->h_state, b->d_state, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, b->stream),
              "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    bool all = true;
    for (int i = 0; i < nb; ++i) all = all && b->h_state[2 * i + 1] != 0.0;
    if (all) break;
    batch = 2;
  }
  for (int i = 0; i < nb; ++i) {
    tpx_ctx *c = b->ctx[i];
    const int it = (int)b->h_state[2 * i];
    c->iter_guess[rk & 3] = it;
    c->num_iteration = it;
    c->total_iterations += it;
  }
}

// ---- gas adjoint: recomputation and the reverse sweep -----------------------
// The gas state of a step boundary, RU, RU_int, Rho (= Rho_int) and P, saved
// to or restored from dst (8 n values).
void gas_state_copy(tpx_ctx *c, double *dst, bool save) {
  const long n = c->g.n;
  if (c->ru_int == c->ru_base || !c->ru_np1_is_base || !c->rho_np1_is_base)
    throw std::logic_error("adjoint: the gas is not at a step boundary");
  double *f[8] = {c->ru[c->ru_base][0], c->ru[c->ru_base][1], c->ru[c->ru_base][2], c->ru[c->ru_int][0],
                  c->ru[c->ru_int][1],  c->ru[c->ru_int][2],  c->rho[c->rho_base], c->P};
  for (int q = 0; q < 8; ++q) {
    if (save) launch_copy(c, dst + q * n, f[q], n);
    else launch_copy(c, f[q], dst + q * n, n);
  }
  if (!save && c->rho_int != c->rho_base) launch_copy(c, c->rho[c->rho_int], dst + 6 * n, n);
}

// One pool buffer of n values (grown on demand); null when out of memory.
double *pcap_slot(AdjointTape &t, size_t i, long n) {
  while (t.pcap.size() <= i) {
    double *p = nullptr;
    if (hipMalloc((void **)&p, sizeof(double) * n) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    t.pcap.push_back(p);
  }
  return t.pcap[i];
}

// The gas substeps of step s of the recorded run, gas only (the gas does not
// depend on the particles under one-way coupling; under two-way coupling the
// recorded drag deposits stand in for them, with the heat exchange the
// recorded RHS_Part_Temp, its sum and P0_new), from the gas state in place
// at the step's start, keeping what the reverse sweep reads: RU_new and
// Rho_new after the momentum and density updates, the raw Poisson right-hand
// side and its sum, and P before every Poisson iteration.  The arithmetic is
// the fused step's.  With `segment` set the step's substep starts (RU_int,
// Rho_int) are taped too, into the loaded segment's slots, and the captures
// go to the segment's sub tape; else to the one step's capture buffers.
void gas_substeps(tpx_ctx *c, int s, GasSubstep sub[4], bool segment) {
  AdjointTape &t = c->adj;
  const long n = c->g.n;
  c->A = t.A[s];
  c->Iteration = t.iteration[s];
  for (int d = 0; d < 3; ++d) c->gv[d] = t.gv[(size_t)s * 3 + d];
  const int j = s - t.gas_base;  // slot in a segment's tapes
  size_t used = segment ? t.pnext : 0;
  for (int rk = 0; rk < 4; ++rk) {
    Coef k = coef(c, rk);
    GasSubstep &g = sub[rk];
    double *gs = t.gas_of(s, n) + (long)rk * 4 * n;
    if (segment) {
      for (int d = 0; d < 3; ++d) launch_copy(c, gs + d * n, c->ru[c->ru_int][d], n);
      launch_copy(c, gs + 3 * n, c->rho[c->rho_int], n);
    }
    for (int d = 0; d < 3; ++d) g.ru_int[d] = gs + d * n;
    g.rho_int = gs + 3 * n;
    const size_t q = (size_t)s * 4 + rk;
    if (t.heat) {  // Update_P0's P0_new, and the substep's deposit for the divergence (read from the tape)
      launch_copy(c, c->scal + SC_P0_NEW, t.p0t + q + 1, 1);
      launch_copy(c, c->scal + SC_SUM_Q, t.qs + q, 1);
      c->q_zero = false;
      g.Q = t.qt + (long)q * n;
      g.q_sum = t.qs + q;
      g.p0n = t.p0t + q + 1;
    } else {
      launch_update_p0(c, k);
    }
    {
      // two-way coupling: the momentum update reads the recorded drag deposit
      // of this substep in place (the context's S is pointed at the tape for
      // the launch; nothing in a gas-only step writes S)
      double *keep[3] = {c->S[0], c->S[1], c->S[2]};
      if (t.S)
        for (int d = 0; d < 3; ++d) c->S[d] = t.S + (((long)s * 4 + rk) * 3 + d) * n;
      try {
        do_update_ru_wop(c, rk, true);
      } catch (...) {
        for (int d = 0; d < 3; ++d) c->S[d] = keep[d];
        throw;
      }
      for (int d = 0; d < 3; ++d) c->S[d] = keep[d];
    }
    double *cp = segment ? t.sub + ((long)j * 4 + rk) * 5 * n : t.cap + (long)rk * 5 * n;
    for (int d = 0; d < 3; ++d) {
      launch_copy(c, cp + d * n, c->ru[c->ru_new][d], n);
      g.ru_new[d] = cp + d * n;
    }
    launch_copy(c, cp + 3 * n, c->rho[c->rho_new], n);
    g.rho_new = cp + 3 * n;
    const double *rn[3];
    ru_ptrs(c, c->ru_new, rn);
    launch_div_rhs(c, k, 0, rn, c->rho[c->rho_new], t.heat ? g.Q : (c->q_zero ? nullptr : c->Q), nullptr, c->rhs);
    launch_copy(c, cp + 4 * n, c->rhs, n);
    g.rhs = cp + 4 * n;
    double *sum = segment ? t.sub + (long)t.tape_steps() * 20 * n + (long)j * 4 + rk : t.cap + 20 * n + rk;
    launch_copy(c, sum, c->scal + SC_SUM_RHS, 1);
    g.rhs_sum = sum;
    // the recorded number of iterations, P kept before each: the first P is
    // copied into its slot, and every iteration writes its result into the
    // next slot (the last into P), out of place, so P before each stays; the
    // count the device's loop test reaches is kept for a check after the
    // sweep (the replay is the recorded run bit for bit, so no iteration is
    // skipped on the device before the recorded count)
    launch_solve_begin(c);
    g.P.clear();
    g.iters = t.iters[q];
    if (segment) t.p_first[q] = used;
    for (int it = 0; it < g.iters; ++it) {
      double *buf = pcap_slot(t, used++, n);
      if (!buf) throw std::bad_alloc();
      if (it == 0) launch_copy(c, buf, c->P, n);
      g.P.push_back(buf);
      double *out = c->P;
      if (it + 1 < g.iters && !(out = pcap_slot(t, used, n))) throw std::bad_alloc();
      launch_poisson_iteration_on(c, c->rho[c->rho_new], buf, c->rhs, c->scal, nullptr, SC_SUM_RHS, out);
    }
    launch_copy(c, t.iters_seen + (long)s * 4 + rk, c->scal + SC_ITERS, 1);
    do_update_ru_wp(c, rk, rk == 3);
    c->rho_int = c->rho_new;  // the gas half of TimeAdvance_RK4
    c->ru_int = c->ru_new;
    c->rho_new = -1;
    c->ru_new = -1;
    launch_scalar_copy(c, SC_P0_INT, SC_P0_NEW);
  }
  if (segment) t.pnext = used;
  const double a[3] = {c->p.u0 * c->p.rho0, c->p.v0 * c->p.rho0, c->p.w0 * c->p.rho0};
  do_make_mean_u0(c, a, true);
  c->rho_base = rho_np1_read(c);  // the gas half of TimeAdvance
  c->rho_np1_is_base = true;
  c->ru_base = ru_np1_read(c);
  c->ru_np1_is_base = true;
  launch_scalar_copy(c, SC_P0, SC_P0_NP1);
}

// Recomputes step s of the recorded run from its start on the tape (RU and P
// of the step's start, RU_int and Rho of its first substep), capturing into
// the one step's buffers.
void gas_recompute_step(tpx_ctx *c, int s, GasSubstep sub[4]) {
  AdjointTape &t = c->adj;
  const long n = c->g.n;
  if (c->ru_int == c->ru_base) throw std::logic_error("adjoint: RU_int aliases RU at a step boundary");
  const double *st = t.start + (long)s * 4 * n;
  const double *g0 = t.gas_of(s, n);
  for (int d = 0; d < 3; ++d) {
    launch_copy(c, c->ru[c->ru_base][d], st + d * n, n);
    launch_copy(c, c->ru[c->ru_int][d], g0 + d * n, n);
  }
  launch_copy(c, c->rho[c->rho_base], g0 + 3 * n, n);
  if (c->rho_int != c->rho_base) launch_copy(c, c->rho[c->rho_int], g0 + 3 * n, n);
  launch_copy(c, c->P, st + 3 * n, n);
  gas_substeps(c, s, sub, false);
}

// Segment mode: replays checkpoint segment k (its steps, gas only, from the
// gas checkpoint of its first step), taping every substep's start and what
// the reverse sweep reads into the segment tapes.
void gas_replay_segment(tpx_ctx *c, int k) {
  AdjointTape &t = c->adj;
  const long n = c->g.n;
  const int base = k * t.every, end = std::min(t.nsteps, base + t.every);
  gas_state_copy(c, t.gck + (long)k * 8 * n, false);
  t.gas_base = base;
  t.pnext = 0;
  GasSubstep sub[4];  // filled per step; the sweep reads the segment tapes through taped_step
  for (int s = base; s < end; ++s) gas_substeps(c, s, sub, true);
  t.seg_loaded = k;
}

// Step s of the recorded run as the gas sweep reads it, from the tape kept
// while recording.
void taped_step(tpx_ctx *c, int s, GasSubstep sub[4]) {
  AdjointTape &t = c->adj;
  const long n = c->g.n;
  c->A = t.A[s];
  c->Iteration = t.iteration[s];
  for (int d = 0; d < 3; ++d) c->gv[d] = t.gv[(size_t)s * 3 + d];
  const int j = s - t.gas_base;  // slot in the tapes (a segment's, or the whole run's)
  for (int rk = 0; rk < 4; ++rk) {
    GasSubstep &g = sub[rk];
    const size_t q = (size_t)s * 4 + rk;
    const double *gs = t.gas_of(s, n) + (long)rk * 4 * n, *sb = t.sub + ((long)j * 4 + rk) * 5 * n;
    for (int d = 0; d < 3; ++d) {
      g.ru_int[d] = gs + d * n;
      g.ru_new[d] = sb + d * n;
    }
    g.rho_int = gs + 3 * n;
    g.rho_new = sb + 3 * n;
    g.rhs = sb + 4 * n;
    g.rhs_sum = t.sub + (long)t.tape_steps() * 20 * n + (long)j * 4 + rk;
    if (t.heat) {
      g.Q = t.qt + (long)q * n;
      g.q_sum = t.qs + q;
      g.p0n = t.p0t + q + 1;
    }
    g.iters = t.iters[q];
    g.P.clear();
    for (int i = 0; i < g.iters; ++i) g.P.push_back(t.pcap[t.p_first[q] + i]);
  }
}

// The reverse sweep of the whole recorded run, particles and gas together:
// after the particles' step s has deposited dJ/du_g of its four substeps,
// the gas step s is recomputed and reversed.  The solver's gas ends where the
// recorded run ended.
void gas_adjoint_run(tpx_ctx *c) {
  AdjointTape &t = c->adj;
  const long n = c->g.n;
  auto alloc = [](double **p, size_t count) {
    if (!*p && hipMalloc((void **)p, sizeof(double) * count) != hipSuccess) throw std::bad_alloc();
  };
  alloc(&t.work, (size_t)GasAdjoint::kFields * n + SC_COUNT);
  alloc(&t.lup, 3 * (size_t)n);
  if (!t.store) alloc(&t.iters_seen, 4 * (size_t)t.nsteps);
  if (!t.store && !t.segmented) alloc(&t.cap, 20 * (size_t)n + 8);
  if (t.segmented) alloc(&t.sub, (size_t)t.tape_steps() * (20 * (size_t)n + 4));
  GasAdjoint &ga = t.ga;
  double *w = t.work;
  auto take = [&]() { double *p = w; w += n; return p; };
  for (auto *grp : {ga.lru, ga.lnp1, ga.lint, ga.lnew, ga.lrhs, ga.lF})
    for (int d = 0; d < 3; ++d) grp[d] = take();
  for (double **p : {&ga.lr, &ga.lrnp1, &ga.lrint, &ga.lrnew, &ga.lp, &ga.lrhs_p, &ga.lcc, &ga.ld, &ga.nu, &ga.cc,
                     &ga.lq})
    *p = take();
  ga.scal = w;
  launch_fill(c, t.work, 0.0, (long)GasAdjoint::kFields * n + SC_COUNT);
  launch_fill(c, c->scal + SC_ADJ_A, 0.0, 1);
  if (t.heat) {
    launch_fill(c, c->scal + SC_ADJ_P0, 0.0, SC_ADJ_P0I - SC_ADJ_P0 + 1);
    // a recomputing sweep rewrites the Update_P0 slots and Q: keep the run's end
    if (!t.store) launch_copy(c, t.scal_save, c->scal, SC_SUM_Q + 1);
  }
  const double A0 = c->A, g0[3] = {c->gv[0], c->gv[1], c->gv[2]};
  const int it0 = c->Iteration;
  GasSubstep sub[4];
  const bool stored = t.store;
  GasSweep sweep;
  sweep.begin = [&](int s) {
    if (stored || t.segmented) taped_step(c, s, sub);
    else gas_recompute_step(c, s, sub);
    if (s < t.nsteps - 1) adjoint_gas_boundary(c, ga, t.heat);
  };
  if (t.segmented) sweep.segment = [&](int k) { gas_replay_segment(c, k); };
  sweep.part_a = [&](int, int rk) { adjoint_gas_substep_a(c, ga, sub[rk], coef(c, rk)); };
  sweep.part_rk = [&](int, int rk) { return adjoint_gas_substep_rk(c, ga, sub[rk], coef(c, rk)); };
  sweep.heat_mid = [&](int, int rk) { return GasSweep::HeatMid{sub[rk].rho_new, ga.lq, ga.lrnew}; };
  sweep.heat_init = GasSweep::HeatMid{t.gas + 3 * n, nullptr, ga.lr};
  sweep.part_b = [&](int, int rk, const double *lup) {
    sub[rk].lup = lup;
    adjoint_gas_substep_b(c, ga, sub[rk], coef(c, rk));
  };
  sweep.a_drag = c->p.twc ? -c->mp / c->Vcell : 0.0;
  // the run's end where a recomputing sweep rewrote it
  auto restore = [&]() {
    c->A = A0; c->Iteration = it0;
    for (int d = 0; d < 3; ++d) c->gv[d] = g0[d];
    if (stored) return;
    gas_state_copy(c, t.final_state, false);
    if (t.heat) {
      launch_copy(c, c->scal, t.scal_save, SC_SUM_Q + 1);
      launch_copy(c, c->Q, t.qt + ((long)t.nsteps * 4 - 1) * n, n);
    }
  };
  try {
    adjoint_reverse(c, &sweep);
  } catch (...) {
    restore();
    throw;
  }
  restore();
  t.swept = true;
  if (stored) return;
  // every recomputed solve stopped where the recorded one did
  std::vector<double> seen((size_t)t.nsteps * 4);
  hip_check(hipMemcpyAsync(seen.data(), t.iters_seen, sizeof(double) * seen.size(), hipMemcpyDeviceToHost, c->stream),
            "download");
  sync(c);
  for (size_t q = 0; q < seen.size(); ++q)
    if ((int)seen[q] != t.iters[q]) throw std::logic_error("adjoint: a recomputed Poisson solve differs from the recorded one");
}

// host copy of the sum of device arrays (n values each)
void download_sum(tpx_ctx *c, std::initializer_list<const double *> src, double *dst) {
  const size_t n = (size_t)c->g.n;
  std::vector<double> tmp(n);
  bool first = true;
  for (const double *p : src) {
    hip_check(hipMemcpyAsync(first ? dst : tmp.data(), p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream),
              "download");
    sync(c);
    if (!first)
      for (size_t q = 0; q < n; ++q) dst[q] += tmp[q];
    first = false;
  }
}

}  // namespace

extern "C" {

const char *tpx_batch_error(void) { return g_batch_error.c_str(); }

int tpx_batch_create(tpx_ctx *const *ctxs, int n, tpx_batch **out) {
  if (!out) return TPX_E_ARG;
  *out = nullptr;
  if (!ctxs || n < 1) {
    g_batch_error = "a batch needs at least one context";
    return TPX_E_ARG;
  }
  std::vector<tpx_ctx *> seen(ctxs, ctxs + n);
  std::sort(seen.begin(), seen.end());
  for (int i = 0; i < n; ++i)
    if (!seen[i] || (i > 0 && seen[i] == seen[i - 1])) {
      g_batch_error = "null or repeated context";
      return TPX_E_ARG;
    }
  const tpx_ctx *a = ctxs[0];
  for (int i = 1; i < n; ++i) {  // the switches that decide which kernels a step launches
    const tpx_ctx *c = ctxs[i];
    const char *why = nullptr;
    if (c->g.nx != a->g.nx || c->g.ny != a->g.ny || c->g.nz != a->g.nz) why = "grid sizes differ";
    else if (c->device != a->device) why = "contexts on different devices";
    else if ((c->p.twc != 0) != (a->p.twc != 0)) why = "TWC differs";
    else if (heat_active(c) != heat_active(a)) why = "heat exchange on in some boxes only";
    else if (c->p.solve_for_scalar) why = "the passive scalar is not batched (use tpx_ensemble_step)";
    else if (c->p.solve_for_vector) why = "the passive vector is not batched (use tpx_ensemble_step)";
    else if ((c->pt.n > 0) != (a->pt.n > 0)) why = "particles in some boxes only";
    else if (c->pt.has_tau != a->pt.has_tau || c->pt.has_eps != a->pt.has_eps) why = "per-particle tau/eps differ";
    else if (c->fft_passes != a->fft_passes) why = "Poisson pass counts differ";
    if (why) {
      g_batch_error = why;
      return TPX_E_ARG;
    }
  }
  tpx_batch *b = new tpx_batch();
  try {
    b->ctx.assign(ctxs, ctxs + n);
    b->rec.resize(n);
    b->device = a->device;
    if (const char *e = getenv("TPX_BATCH_GRAPH")) b->graphs = atoi(e) != 0;
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto &e : b->ev) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    if (hipMalloc((void **)&b->d_state, sizeof(double) * 2 * n) != hipSuccess) throw std::bad_alloc();
    hip_check(hipHostMalloc((void **)&b->h_state, sizeof(double) * 2 * n), "hipHostMalloc");
  } catch (const std::exception &e) {
    g_batch_error = e.what();
    batch_release(b);
    return TPX_E_HIP;
  }
  *out = b;
  return TPX_OK;
}

int tpx_batch_step(tpx_batch *b, int nsteps) {
  if (!b) return TPX_E_ARG;
  if (nsteps < 0) {
    b->err = "nsteps < 0";
    return TPX_E_ARG;
  }
  try {
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    for (tpx_ctx *c : b->ctx) {  // work queued on the boxes' own streams comes first
      if (c->tm.on) throw std::logic_error("kernel timing is not recorded in a batch");
      if (c->tile.active) throw std::logic_error("a box is still tuning its tile shapes: step it alone first");
      sync(c);
    }
    for (int s = 0; s < nsteps; ++s) {
      for (tpx_ctx *c : b->ctx) do_params_update(c);
      for (int rk = 0; rk < 4; ++rk) {
        batch_run_graph(b, [rk](tpx_ctx *c, int) { fused_substep_pre(c, rk); });
        batch_solve(b, rk);
        batch_run_graph(b, [rk](tpx_ctx *c, int) { fused_substep_post(c, rk); });
      }
      batch_run_graph(b, [](tpx_ctx *c, int) { step_end(c); });
    }
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  } catch (const std::logic_error &e) {
    b->err = e.what();
    return TPX_E_STATE;
  } catch (const std::bad_alloc &) {
    b->err = "out of memory";
    return TPX_E_OOM;
  } catch (const std::exception &e) {
    b->err = e.what();
    return TPX_E_HIP;
  }
  return TPX_OK;
}

const char *tpx_batch_last_error(tpx_batch *b) { return b ? b->err.c_str() : "null batch"; }

int tpx_batch_destroy(tpx_batch *b) {
  if (!b) return TPX_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_release(b);
  return TPX_OK;
}

// ---- adjoint of the particle path ----------------------------------------
int tpx_adjoint_record(tpx_ctx *ctx, int nsteps, int every, double *J) {
  return guarded(ctx, [&]() {
    tpx_ctx *c = ctx;
    Particles &pt = c->pt;
    if (nsteps < 1 || every < 1) throw std::invalid_argument("nsteps and every must be positive");
    if (pt.n == 0) throw std::invalid_argument("the adjoint needs particles");
    AdjointTape &t = c->adj;
    sync(c);
    adjoint_free(t);
    t.nsteps = nsteps;
    t.every = every;
    t.heat = heat_active(c);
    t.nq = t.heat ? 7 : 6;
    const long n = c->g.n, np = pt.n;
    const size_t nq = (size_t)t.nq;
    auto alloc = [](double **p, size_t count) {
      if (hipMalloc((void **)p, sizeof(double) * count) != hipSuccess) throw std::bad_alloc();
    };
    t.segmented = t.mode == 3;
    t.gas_base = 0;
    t.seg_loaded = -1;
    alloc(&t.gas, (size_t)t.tape_steps() * 16 * n);
    alloc(&t.ckpt, (size_t)(nsteps / every + 1) * nq * np);
    if (t.segmented) alloc(&t.gck, (size_t)((nsteps - 1) / every + 1) * 8 * n);
    alloc(&t.y, nq * (size_t)np);
    alloc(&t.stage, 4 * nq * (size_t)np);
    alloc(&t.lam, nq * (size_t)np);
    if (t.heat) {  // RHS_Part_Temp, its sum and P0 of every substep
      alloc(&t.qt, (size_t)nsteps * 4 * n);
      alloc(&t.qs, (size_t)nsteps * 4);
      alloc(&t.p0t, (size_t)nsteps * 4 + 1);
      alloc(&t.deps, (size_t)np);
      alloc(&t.lh, (size_t)np);
      alloc(&t.hpart, (size_t)grid1d(np, 256));
      alloc(&t.scal_save, SC_COUNT);
      if (pt.has_eps) {
        alloc(&t.eps0, (size_t)np);
        launch_copy(c, t.eps0, pt.eps, np);
      }
      t.I0.assign((size_t)nsteps, 0.0);
    }
    alloc(&t.dtau, (size_t)np);
    alloc(&t.djdc, (size_t)n);
    if (hipMalloc((void **)&t.cum, sizeof(int32_t) * np) != hipSuccess ||
        hipMalloc((void **)&t.cum_tmp, sizeof(int32_t) * np) != hipSuccess)
      throw std::bad_alloc();
    launch_iota(c, t.cum, np);
    if (pt.has_tau) {
      alloc(&t.tau0, (size_t)np);
      launch_copy(c, t.tau0, pt.tau, np);
    }
    if (!t.segmented) alloc(&t.start, (size_t)nsteps * 4 * n);
    const bool twc = c->p.twc != 0;
    if (twc) alloc(&t.S, (size_t)nsteps * 12 * n);  // the drag deposits, for a gas recomputation
    alloc(&t.final_state, 8 * (size_t)n);
    // keep every substep's momentum, density and Poisson data when the
    // estimate (three Poisson iterations a solve) fits in half the free HBM
    t.p_first.assign((size_t)nsteps * 4, 0);
    t.pnext = 0;
    t.store = false;
    if (t.mode == 0 || t.mode == 1) {
      size_t free_b = 0, total_b = 0;
      hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
      const double need = (double)nsteps * 4 * (5 + 3) * n * sizeof(double);
      if (t.mode == 1 || need < 0.5 * (double)free_b) {
        t.store = hipMalloc((void **)&t.sub, sizeof(double) * ((size_t)nsteps * 20 * n + (size_t)nsteps * 4)) ==
                  hipSuccess;
        if (!t.store) {
          (void)hipGetLastError();
          t.sub = nullptr;
          if (t.mode == 1) throw std::bad_alloc();
        }
      }
    }
    t.g3.assign((size_t)nsteps * 3, 0.0);
    t.A.assign((size_t)nsteps, 0.0);
    t.gv.assign((size_t)nsteps * 3, 0.0);
    t.iteration.assign((size_t)nsteps, 0);
    t.iters.assign((size_t)nsteps * 4, 0);
    adjoint_checkpoint(c, 0);
    try {
      for (int s = 0; s < nsteps; ++s) {
        t.step = s;
        do_params_update(c);
        if (t.segmented) {
          // the gas at a segment's start: RU, RU_int, Rho, P (at the run's
          // start RU_int is RU's own buffer; the restore writes both)
          if (s % every == 0) {
            double *dst = t.gck + (long)(s / every) * 8 * n;
            for (int d = 0; d < 3; ++d) {
              launch_copy(c, dst + d * n, c->ru[c->ru_base][d], n);
              launch_copy(c, dst + (3 + d) * n, c->ru[c->ru_int][d], n);
            }
            launch_copy(c, dst + 6 * n, c->rho[c->rho_base], n);
            launch_copy(c, dst + 7 * n, c->P, n);
          }
        } else {
          // the gas at the step's start (RU_int and Rho are the first substep's tape)
          for (int d = 0; d < 3; ++d) launch_copy(c, t.start + ((long)s * 4 + d) * n, c->ru[c->ru_base][d], n);
          launch_copy(c, t.start + ((long)s * 4 + 3) * n, c->P, n);
        }
        t.A[s] = c->A;
        if (t.heat) t.I0[s] = c->I0;
        t.iteration[s] = c->Iteration;
        for (int d = 0; d < 3; ++d) t.gv[(size_t)s * 3 + d] = c->gv[d];
        for (int rk = 0; rk < 4; ++rk) {  // fused_substep, keeping each solve's iteration count
          fused_substep_pre(c, rk);
          const size_t q = (size_t)s * 4 + rk;
          if (twc)
            for (int d = 0; d < 3; ++d) launch_copy(c, t.S + ((long)q * 3 + d) * n, c->S[d], n);
          if (t.heat) {
            launch_copy(c, t.qt + (long)q * n, c->Q, n);
            launch_copy(c, t.qs + q, c->scal + SC_SUM_Q, 1);
          }
          if (t.store) {  // what the gas sweep reads, so that it need not recompute the step
            double *dst = t.sub + (long)q * 5 * n;
            for (int d = 0; d < 3; ++d) launch_copy(c, dst + d * n, c->ru[c->ru_new][d], n);
            launch_copy(c, dst + 3 * n, c->rho[c->rho_new], n);
            launch_copy(c, dst + 4 * n, c->rhs, n);
            launch_copy(c, t.sub + (long)nsteps * 20 * n + q, c->scal + SC_SUM_RHS, 1);
          }
          const size_t p0 = t.pnext;
          t.p_first[q] = p0;
          auto keep_p = [&](int i) {
            if (!t.store) return;
            double *buf = pcap_slot(t, p0 + (size_t)i, n);
            if (!buf) {  // out of memory: the sweep will recompute instead
              t.store = false;
              return;
            }
            launch_copy(c, buf, c->P, n);
          };
          do_solve_poisson(c, rk, &t.iters[q], keep_p);
          t.pnext = p0 + (size_t)t.iters[q];
          fused_substep_post(c, rk);
        }
        step_end(c);
        t.step = -1;
        if ((s + 1) % every == 0 && s + 1 < nsteps) adjoint_checkpoint(c, (s + 1) / every);
      }
    } catch (...) {
      t.step = -1;
      throw;
    }
    gas_state_copy(c, t.final_state, true);
    if (t.heat) launch_copy(c, t.p0t + (long)nsteps * 4, c->scal + SC_P0, 1);
    const double j = adjoint_objective(c);
    if (J) *J = j;
    t.ready = true;
    return TPX_OK;
  });
}

int tpx_adjoint_gradient(tpx_ctx *ctx, double *dJ_dtau, double *dJ_dstate) {
  return tpx_adjoint_gradient_full(ctx, dJ_dtau, dJ_dstate, nullptr, nullptr, nullptr, nullptr);
}

int tpx_adjoint_gradient_full(tpx_ctx *ctx, double *dJ_dtau, double *dJ_dstate, double *dJ_dA, double *dJ_dRU,
                              double *dJ_dRho, double *dJ_dP) {
  return guarded(ctx, [&]() {
    AdjointTape &t = ctx->adj;
    if (!t.ready) throw std::logic_error("no recorded run: call tpx_adjoint_record first");
    const bool gas = dJ_dA || dJ_dRU || dJ_dRho || dJ_dP;
    if (!gas && !dJ_dtau && !dJ_dstate) throw std::invalid_argument("null outputs");
    // under two-way coupling or with the heat exchange the gas depends on
    // the particles, so every gradient needs the gas sweep
    // the same holds in the segment mode, whose sweep replays the gas tape
    if (gas || ctx->p.twc || t.heat || t.segmented) gas_adjoint_run(ctx);
    else adjoint_reverse(ctx, nullptr);
    const size_t np = (size_t)ctx->pt.n, n = (size_t)ctx->g.n;
    // the sweep's particle arrays are in the recording's order: give them in
    // the context's present order (cum maps the one to the other)
    std::vector<int32_t> cum(np);
    std::vector<double> tmp(6 * np);
    hip_check(hipMemcpyAsync(cum.data(), t.cum, sizeof(int32_t) * np, hipMemcpyDeviceToHost, ctx->stream), "download");
    if (dJ_dtau)
      hip_check(hipMemcpyAsync(tmp.data(), t.dtau, sizeof(double) * np, hipMemcpyDeviceToHost, ctx->stream), "download");
    sync(ctx);
    if (dJ_dtau)
      for (size_t i = 0; i < np; ++i) dJ_dtau[i] = tmp[(size_t)cum[i]];
    if (dJ_dstate) {  // adjoint of the recorded run's initial x, y, z, u, v, w (the first 6 rows)
      hip_check(hipMemcpyAsync(tmp.data(), t.lam, sizeof(double) * 6 * np, hipMemcpyDeviceToHost, ctx->stream),
                "download");
      sync(ctx);
      for (int q = 0; q < 6; ++q)
        for (size_t i = 0; i < np; ++i) dJ_dstate[q * np + i] = tmp[q * np + (size_t)cum[i]];
    }
    if (dJ_dA)
      hip_check(hipMemcpyAsync(dJ_dA, ctx->scal + SC_ADJ_A, sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                "download");
    sync(ctx);
    const GasAdjoint &ga = t.ga;
    // the recorded run starts with RU_int = RU_np1 = RU and Rho_int = Rho_np1 = Rho
    if (dJ_dRU)
      for (int d = 0; d < 3; ++d) download_sum(ctx, {ga.lru[d], ga.lnp1[d], ga.lint[d]}, dJ_dRU + d * n);
    if (dJ_dRho) download_sum(ctx, {ga.lr, ga.lrnp1, ga.lrint}, dJ_dRho);
    if (dJ_dP) download_sum(ctx, {ga.lp}, dJ_dP);
    return TPX_OK;
  });
}

int tpx_adjoint_gradient_heat(tpx_ctx *ctx, double *dJ_dT, double *dJ_deps, double *dJ_dP0) {
  return guarded(ctx, [&]() {
    AdjointTape &t = ctx->adj;
    if (!t.ready) throw std::logic_error("no recorded run: call tpx_adjoint_record first");
    if (!t.heat) throw std::logic_error("the recorded run has no heat exchange");
    if (!dJ_dT && !dJ_deps && !dJ_dP0) throw std::invalid_argument("null outputs");
    if (!t.swept) gas_adjoint_run(ctx);
    const size_t np = (size_t)ctx->pt.n;
    std::vector<int32_t> cum(np);
    std::vector<double> tmp(np);
    hip_check(hipMemcpyAsync(cum.data(), t.cum, sizeof(int32_t) * np, hipMemcpyDeviceToHost, ctx->stream), "download");
    sync(ctx);
    auto give = [&](const double *src, double *dst) {  // in the context's present particle order
      hip_check(hipMemcpyAsync(tmp.data(), src, sizeof(double) * np, hipMemcpyDeviceToHost, ctx->stream), "download");
      sync(ctx);
      for (size_t i = 0; i < np; ++i) dst[i] = tmp[(size_t)cum[i]];
    };
    if (dJ_dT) give(t.lam + 6 * np, dJ_dT);
    if (dJ_deps) give(t.deps, dJ_deps);
    if (dJ_dP0) {  // the initial P0 is both P0 and P0_np1 of the first step, and fed its first exchange
      double sc[SC_ADJ_P0I - SC_ADJ_P0 + 1];
      hip_check(hipMemcpyAsync(sc, ctx->scal + SC_ADJ_P0, sizeof(sc), hipMemcpyDeviceToHost, ctx->stream), "download");
      sync(ctx);
      *dJ_dP0 = sc[0] + sc[SC_ADJ_P0NP1 - SC_ADJ_P0] + sc[SC_ADJ_P0I - SC_ADJ_P0];
    }
    return TPX_OK;
  });
}

int tpx_adjoint_set_gas_tape(tpx_ctx *ctx, int mode) {
  return guarded(ctx, [&]() {
    if (mode < 0 || mode > 3)
      throw std::invalid_argument("gas tape mode must be 0 (by memory), 1 (keep), 2 (recompute) or 3 (segments)");
    ctx->adj.mode = mode;
    return TPX_OK;
  });
}

int tpx_adjoint_release(tpx_ctx *ctx) {
  return guarded(ctx, [&]() {
    sync(ctx);
    adjoint_free(ctx->adj);
    return TPX_OK;
  });
}

int tpx_adjoint_check(tpx_ctx *ctx, int op, const double *x, const double *y, double *Ax, double *ATy,
                      double *ATy_rho) {
  return guarded(ctx, [&]() {
    if (op < 0 || op > 1) throw std::invalid_argument("op must be 0 (spectral step) or 1 (Poisson iteration)");
    if (!x || !y || !Ax || !ATy || (op == 1 && !ATy_rho)) throw std::invalid_argument("null host array");
    const size_t n = (size_t)ctx->g.n;
    sync(ctx);
    double *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(double) * (13 * n + SC_COUNT)) != hipSuccess) throw std::bad_alloc();
    try {
      hip_check(hipMemcpyAsync(d, x, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream), "upload");
      hip_check(hipMemcpyAsync(d + n, y, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream), "upload");
      adjoint_poisson_check(ctx, op, ctx->rho[ctx->rho_base], d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n);
      hip_check(hipMemcpyAsync(Ax, d + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream), "download");
      hip_check(hipMemcpyAsync(ATy, d + 3 * n, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream), "download");
      if (op == 1)
        hip_check(hipMemcpyAsync(ATy_rho, d + 4 * n, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream),
                  "download");
      sync(ctx);
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(d);
      throw;
    }
    (void)hipFree(d);
    return TPX_OK;
  });
}

int tpx_field_sum(tpx_ctx *ctx, int field, int comp, double *out) {
  return guarded(ctx, [&]() {
    const double *f = field_buffer(ctx, field, comp);
    if (!f || !out) throw std::invalid_argument("bad field");
    launch_reduce(ctx, f, 0, SC_RED);
    *out = read_scalar(ctx, SC_RED);
    return TPX_OK;
  });
}

int tpx_field_max_abs(tpx_ctx *ctx, int field, int comp, double *out) {
  return guarded(ctx, [&]() {
    const double *f = field_buffer(ctx, field, comp);
    if (!f || !out) throw std::invalid_argument("bad field");
    launch_reduce(ctx, f, 1, SC_RED);
    *out = read_scalar(ctx, SC_RED);
    return TPX_OK;
  });
}

int tpx_deposit_concentration(tpx_ctx *ctx) {
  return guarded(ctx, [&]() {
    if (ctx->pt.n == 0) launch_fill(ctx, ctx->C, 0.0, ctx->g.n);
    else launch_concentration(ctx);
    sync(ctx);
    return TPX_OK;
  });
}

int tpx_get_stream(tpx_ctx *ctx, void **stream) {
  return guarded(ctx, [&]() {
    if (!stream) throw std::invalid_argument("null");
    *stream = (void *)ctx->stream;
    return TPX_OK;
  });
}

// Deposit tile shapes: tuned online inside the next steps, or set
int tpx_tune_tiles(tpx_ctx *ctx) {
  return guarded(ctx, [&]() {
    tile_tune_begin(ctx);
    return TPX_OK;
  });
}

int tpx_get_tiles(tpx_ctx *ctx, int *out) {
  return guarded(ctx, [&]() {
    if (!out) throw std::invalid_argument("null output");
    const TileTune &t = ctx->tile;
    out[0] = t.heat.rows; out[1] = t.heat.zc; out[2] = t.vel.rows; out[3] = t.vel.zc; out[4] = t.active ? 1 : 0;
    return TPX_OK;
  });
}

int tpx_set_tiles(tpx_ctx *ctx, const int *in) {
  return guarded(ctx, [&]() {
    if (!in) throw std::invalid_argument("null input");
    for (int k = 0; k < 2; ++k) {
      const int rows = in[2 * k], zc = in[2 * k + 1];
      const bool builtin = rows == 0 && zc == 0;
      if (!builtin && !tile_shape_valid(ctx, k, rows, zc))
        throw std::invalid_argument("tile shape does not tile the grid or fit in LDS");
    }
    TileTune &t = ctx->tile;
    t.heat = {in[0], in[1]};
    t.vel = {in[2], in[3]};
    t.active = false;
    return TPX_OK;
  });
}

// grid::Statistics (grid.cpp:902-982): reductions on the device at a step
// boundary; the fourth substep's quantities were captured during the step
int tpx_statistics_enable(tpx_ctx *ctx, int on) {
  return guarded(ctx, [&]() {
    ctx->stats_on = on != 0;
    return TPX_OK;
  });
}

int tpx_statistics(tpx_ctx *ctx, double *out) {
  return guarded(ctx, [&]() {
    tpx_ctx *c = ctx;
    if (!out) throw std::invalid_argument("null output");
    if (!c->stats_on) throw std::logic_error("statistics are off: call tpx_statistics_enable first");
    if (!c->ru_np1_is_base || !c->rho_np1_is_base) throw std::logic_error("statistics are taken at a step boundary");
    const double *ru[3];
    ru_ptrs(c, c->ru_base, ru);
    double raw[16];
    stats_reduce(c, c->rho[c->rho_base], ru, raw);
    double sc[SC_COUNT];
    hip_check(hipMemcpy(sc, c->scal, sizeof(sc), hipMemcpyDeviceToHost), "download");
    const double n = (double)c->g.n, dt = c->p.dt;
    for (int q = 0; q < TPX_ST_COUNT; ++q) out[q] = 0.0;
    out[TPX_ST_RHO_MEAN] = raw[0] / n;
    out[TPX_ST_TG_MEAN] = raw[1] / n;
    out[TPX_ST_TKE_U] = raw[2] / n;
    out[TPX_ST_TKE_V] = raw[3] / n;
    out[TPX_ST_TKE_W] = raw[4] / n;
    out[TPX_ST_C_MEAN] = raw[5] / n;
    out[TPX_ST_RHO_MAX] = raw[6];
    out[TPX_ST_RHO_MIN] = raw[7];
    out[TPX_ST_C_MAX] = raw[8];
    out[TPX_ST_C_MIN] = raw[9];
    // U.max() and U.max_cfl(dt) (tensor1.cpp:327-335, 365-373)
    const double d[3] = {c->g.dx, c->g.dy, c->g.dz};
    double umax = raw[10], cfl = raw[10] * dt / d[0];
    for (int a = 1; a < 3; ++a) {
      umax = raw[10 + a] > umax ? raw[10 + a] : umax;
      const double ca = raw[10 + a] * dt / d[a];
      cfl = ca > cfl ? ca : cfl;
    }
    out[TPX_ST_U_MAX] = umax;
    out[TPX_ST_GAS_MAX_CFL] = cfl;
    out[TPX_ST_VP_MAX] = raw[13];
    out[TPX_ST_TP_MAX] = raw[14];
    const double np = (double)c->pt.n;
    out[TPX_ST_TP_MEAN] = c->pt.n > 0 ? raw[15] / np : 0.0;
    out[TPX_ST_HT_MEAN] = sc[SC_MET];
    out[TPX_ST_TKE2] = sc[SC_ST_TKE2] / n;
    out[TPX_ST_UG_MAX] = sc[SC_ST_UG];
    if (c->ps[0]) {  // Passive_Scalar.mean_squares() (grid.cpp:913)
      launch_reduce(c, c->ps[c->ps_base], 2, SC_RED);
      out[TPX_ST_PASSIVE_SCALAR_MEAN] = read_scalar(c, SC_RED) / n;
    }
    if (c->p.solve_for_vector && c->Qv) {
      // V = RV_np1 / Rho_forV; V *= RV_np1; TKEV_V? = V.?.mean() (grid.cpp:918-933),
      // as sums of RV^2 / Rho_forV; VV = V.mean_squares() of the last
      // Compute_RHS_Pois_Q's V (grid.cpp:912)
      const double tkev[3] = {TPX_ST_TKEV_V1, TPX_ST_TKEV_V2, TPX_ST_TKEV_V3};
      for (int a = 0; a < 3; ++a) {
        launch_reduce(c, c->rv[c->rv_base][a], 2, SC_RED);
        out[(int)tkev[a]] = read_scalar(c, SC_RED) / c->p.rho_forv / n;
      }
      out[TPX_ST_VV] = sc[SC_ST_VV] / n;
    }
    out[TPX_ST_P0] = sc[SC_P0];
    out[TPX_ST_PARTICLE_MAX_CFL] = 0.0;
    out[TPX_ST_GAS_MAX_DIFF_CFL] = dt / (c->g.dx * c->g.dx * out[TPX_ST_RHO_MIN] / c->p.mu0) * 6.;
    out[TPX_ST_NUM_ITERATION] = (double)c->num_iteration;
    const double n0 = (double)(int)(c->p.np0 * c->p.lx * c->p.ly * c->p.lz);  // params.h:120
    out[TPX_ST_BALANCE_INDEX] = n0 > 0 ? np / n0 : 0.0;
    return TPX_OK;
  });
}

int tpx_track_particles(tpx_ctx *ctx, int stride, int count, double *xyz) {
  return guarded(ctx, [&]() {
    if (stride <= 0 || count < 0 || (count > 0 && !xyz)) throw std::invalid_argument("bad stride, count or array");
    if (count == 0) return TPX_OK;
    if (ctx->track_cap < count) {
      if (ctx->track) (void)hipFree(ctx->track);
      ctx->track = nullptr;
      ctx->track_cap = 0;
      if (hipMalloc((void **)&ctx->track, sizeof(double) * 3 * (size_t)count) != hipSuccess) throw std::bad_alloc();
      ctx->track_cap = count;
    }
    launch_track(ctx, stride, count, ctx->track);
    hip_check(hipMemcpyAsync(xyz, ctx->track, sizeof(double) * 3 * (size_t)count, hipMemcpyDeviceToHost, ctx->stream),
              "download");
    sync(ctx);
    return TPX_OK;
  });
}

int tpx_timing_enable(tpx_ctx *ctx, int on) {
  return guarded(ctx, [&]() {
    sync(ctx);
    ctx->tm.on = on != 0;
    for (int k = 0; k < TPX_K_COUNT; ++k) { ctx->tm.ms[k] = 0; ctx->tm.launches[k] = 0; }
    for (auto &r : ctx->tm.recs) { ctx->tm.pool.push_back(r.a); ctx->tm.pool.push_back(r.b); }
    ctx->tm.recs.clear();
    return TPX_OK;
  });
}

int tpx_timing_read(tpx_ctx *ctx, double *ms, int64_t *launches) {
  return guarded(ctx, [&]() {
    sync(ctx);
    Timing &t = ctx->tm;
    for (auto &r : t.recs) {
      float e = 0;
      hip_check(hipEventElapsedTime(&e, r.a, r.b), "hipEventElapsedTime");
      t.ms[r.cls] += e;
      t.launches[r.cls] += 1;
      t.pool.push_back(r.a);
      t.pool.push_back(r.b);
    }
    t.recs.clear();
    for (int k = 0; k < TPX_K_COUNT; ++k) {
      if (ms) ms[k] = t.ms[k];
      if (launches) launches[k] = t.launches[k];
    }
    return TPX_OK;
  });
}

int tpx_synchronize(tpx_ctx *ctx) {
  return guarded(ctx, [&]() { sync(ctx); return TPX_OK; });
}

}  // extern "C"
