// tpx_internal.h -- device state and kernel launchers of libtpx (gfx950).
//
// Device layout: every grid field is a dense nz*ny*nx FP64 array with no
// ghost cells, x fastest.  Periodic neighbours are formed inside the kernels
// by index wrap, which replaces the reference's ghost fill after every
// stencil operation (communicator.cpp:108-216).  Vector fields are three such
// arrays.  The RK4 stage copies of the reference (grid.cpp:461-479, 595-606,
// 890-900) are pointer rotations between four buffer sets per field.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/tpx.h"

#define TPX_PI 3.141592653589793  // params.h:5

namespace tpx {

// device scalar slots (ctx->scal)
enum Scal {
  SC_P0 = 0, SC_P0_INT, SC_P0_NEW, SC_P0_NP1, SC_DP0_DT, SC_MET,  // Update_P0 state
  SC_SUM_Q,        // sum of RHS_Part_Temp (interior)
  SC_SUM_RHS,      // sum of the raw Poisson right-hand side
  SC_MAX_DP,       // max |dp| of the last Poisson iteration
  SC_MEAN_RU0, SC_MEAN_RU1, SC_MEAN_RU2,
  SC_RED,          // generic reduction result
  SC_SUM_C,
  SC_ITERS,        // iterations run by the current Poisson solve
  SC_DONE,         // 1: the current solve has stopped (its loop test failed)
  SC_ADJ_SUM,      // adjoint sweep: sum of the Poisson right-hand side's adjoint
  SC_ADJ_M0, SC_ADJ_M1, SC_ADJ_M2,  // adjoint sweep: sums for make_mean_U0's transpose
  SC_ADJ_A,        // adjoint sweep: dJ/dA accumulated
  SC_ST_TKE2,      // statistics: sum of squares of the last substep's pre-projection U
  SC_ST_UG,        // statistics: max gathered gas x velocity of the last substep
  // adjoint sweep with heat exchange: adjoints of the thermodynamic pressure
  // P0 (base, np1 and new of the substep being reversed), of the pending sum
  // of RHS_Part_Temp (the one the next Update_P0 reads), and of the recorded
  // run's initial P0
  SC_ADJ_P0, SC_ADJ_P0NP1, SC_ADJ_P0NEW, SC_ADJ_LSQ, SC_ADJ_P0I,
  SC_PS_S1,        // passive scalar: sum over the cells of its source C_Source
  SC_SUM_RHS_Q,    // passive vector: sum of the raw right-hand side of its Q solve
  SC_RV_S2,        // passive vector: sum over the cells of its source V_Source (y)
  SC_ST_VV,        // statistics: sum of squares of the last substep's V = RV_new / Rho_forV
  SC_COUNT = 32
};

// everything a grid kernel needs, passed by value
struct Geo {
  int nx, ny, nz;
  long nxy, n;
  double dx, dy, dz, idx, idy, idz, idx2, idy2, idz2;
};

struct Coef {             // per-substep constants of the RK4 update
  double dt, pre, post;
  double a_drag, b_visc, c_visc;  // -mp/Vcell, Mu0/3., Mu0
  double A, rho0, g[3];
  double k_over_cp, R, cp, Vcell, cv;
  int rk, cooling;
};

struct Particles {
  int64_t n = 0, cap = 0;
  // q[quantity][array] with quantity x,y,z,u,v,w,T; arrays are four stage
  // buffers whose roles rotate (see st_*)
  double *q[7][4] = {};
  int st_base = 0, st_int = 1, st_new = 2, st_np1 = 3;
  bool np1_is_base = false;        // np1 equals base (start of a step)
  // At the start of a step the int stage equals the base stage; it is then
  // read from the base buffer and st_int names a spare buffer, which the
  // first TimeAdvance_RK4 swap hands to the new stage.
  bool int_is_base = false;
  int int_buf() const { return int_is_base ? st_base : st_int; }
  double *rhs_t = nullptr, *tau = nullptr, *eps = nullptr;
  int32_t *flag = nullptr;
  double *contrib = nullptr;      // 6 per sorted position: velocity deposit shares at x_int
  int32_t *key = nullptr;         // floor cell of x_new (after the wrap), by particle
  int32_t *perm_int = nullptr, *perm_new = nullptr;     // sorted position -> particle
  int32_t *inv_int = nullptr, *inv_new = nullptr;       // particle -> sorted position
  int32_t *start_int = nullptr, *start_new = nullptr;   // CSR over cells (n_cell+1)
  int32_t *cursor = nullptr;      // placement cursors over cells (n_cell+1)
  bool new_counted = false;       // the last advance counted x_new into start_new
  bool has_tau = false, has_eps = false;
  // spare arrays the step-end permutation writes into (rhs_t, tau, eps, flag)
  double *spare[3] = {};
  int32_t *itmp = nullptr;
};

// Work arrays of the gas adjoint sweep (n values each): adjoints of the
// base, np1, int and new momentum and density, of P, of the Poisson
// right-hand side, of the momentum right-hand side and of the face fields F
// (Rho_face) and C (1/Rho_face); the substep's C and cc; the spectral
// solve's buffers.
struct GasAdjoint {
  double *lru[3], *lnp1[3], *lint[3], *lnew[3], *lrhs[3], *lF[3];
  double *lr, *lrnp1, *lrint, *lrnew, *lp, *lrhs_p, *lcc, *ld, *nu, *cc;
  double *lq;    // adjoint of the substep's heat deposit RHS_Part_Temp
  double *scal;  // SC_COUNT slots of the spectral solve (SC_DONE stays 0)
  static constexpr int kFields = 29;
};

// One recomputed substep as the gas adjoint reads it.
struct GasSubstep {
  const double *ru_int[3], *rho_int;   // its start (the tape)
  const double *ru_new[3], *rho_new;   // after the momentum and density updates
  const double *rhs, *rhs_sum;         // raw Poisson right-hand side and its sum (device)
  std::vector<const double *> P;       // P before each Poisson iteration
  int iters = 0;
  const double *lup = nullptr;         // the particles' gather adjoint on the faces (3 n), or null
  // heat exchange (else null): the substep's deposit RHS_Part_Temp, its sum
  // and P0_new (device)
  const double *Q = nullptr, *q_sum = nullptr, *p0n = nullptr;
};

// The adjoint of a recorded run: the gas stage fields of every substep, each
// step's gas start (RU, P), the particle base state every `every` steps, and
// the reverse sweeps' work arrays.  Particle arrays hold nq quantities: x, y,
// z, u, v, w, and T when the heat exchange is on.  In the segment mode the
// gas is checkpointed every `every` steps too, and the per-substep gas tape
// covers only the segment being reversed, replayed from its checkpoint.
struct AdjointTape {
  int nsteps = 0, every = 1;
  int nq = 6;                    // particle quantities taped and reversed (7 with the heat exchange)
  int step = -1;                 // step being recorded (-1: not recording)
  bool ready = false;            // a whole run is recorded and its objective computed
  double *gas = nullptr;         // steps x 4 substeps x (RU_int 3, Rho_int) x n: every step, or one segment's
  int gas_base = 0;              // the step that gas slot 0 holds (a segment's first step, else 0)
  bool segmented = false;        // gas checkpointed every `every` steps, each segment replayed in the sweep
  double *gck = nullptr;         // segment mode: (RU 3, RU_int 3, Rho, P) x n at each segment's start
  int seg_loaded = -1;           // segment mode: the segment whose gas tape is loaded
  std::vector<double> g3;        // particle gravity of each step (3 per step)
  double *ckpt = nullptr;        // (nsteps / every + 1) x nq x np
  double *y = nullptr;           // nq x np
  double *seg = nullptr;         // (min(every, nsteps) + 1) x nq x np: step starts of the segment being reversed
  double *segstage = nullptr;    // min(every, nsteps) x 4 x nq x np: their stage states (null: recomputed per step)
  bool segstage_tried = false;
  double *lb = nullptr, *lnext = nullptr;  // nq x np each: the staged sweep's base and next-stage adjoints
  double *S = nullptr;           // two-way coupling: the drag deposit of every substep (nsteps x 4 x 3 n)
  // particle order: with re-binning (two-way coupling) the particle arrays are
  // re-sorted at every step's end; the tape and the sweep keep the order the
  // recording started in.  cum[i]: that position of the particle now at i.
  int32_t *cum = nullptr, *cum_tmp = nullptr;
  double *tau0 = nullptr;        // per-particle tau in that order (null: scalar tau)
  double *stage = nullptr;       // 4 x nq x np
  double *lam = nullptr;         // nq x np
  double *dtau = nullptr;        // np
  double *djdc = nullptr;        // n
  // gas adjoint
  double *start = nullptr;       // nsteps x (RU 3, P) x n at each step's start
  std::vector<double> A, gv;     // forcing coefficient and gravity of each step (gv: 3 per step)
  std::vector<int> iteration;    // Poisson iteration cap of each step
  std::vector<int> iters;        // iterations each substep's solve ran (4 per step)
  double *iters_seen = nullptr;  // the recomputed solves' counts (4 per step, device)
  double *final_state = nullptr; // RU 3, RU_int 3, Rho, P of the recorded run's end
  double *work = nullptr;        // GasAdjoint::kFields x n + SC_COUNT
  double *cap = nullptr;         // 4 substeps x (RU_new 3, Rho_new, rhs) x n + 4 sums
  double *lup = nullptr;         // 4 substeps x 3 x n
  std::vector<double *> pcap;    // P before each Poisson iteration (grown on demand)
  int mode = 0;                  // gas tape: 0 keep when it fits in memory, 1 keep, 2 recompute each
                                 // step, 3 checkpoint every `every` steps and replay each segment
  bool store = false;            // the recorded run kept every substep (sub, pcap): no recomputation
  double *sub = nullptr;         // steps x 4 x (RU_new 3, Rho_new, raw RHS_Pois) x n, then steps x 4 sums
                                 // (every step, or the loaded segment's)
  std::vector<size_t> p_first;   // first pcap buffer of each substep's solve
  size_t pnext = 0;
  // heat exchange
  bool heat = false;
  double *qt = nullptr;          // nsteps x 4 x n: RHS_Part_Temp of every substep
  double *qs = nullptr;          // nsteps x 4: its sum
  double *p0t = nullptr;         // nsteps x 4 + 1: P0_int at every substep's start, then the final P0
  double *eps0 = nullptr;        // per-particle absorption efficiency in the recording's order (or null)
  double *deps = nullptr;        // np: dJ/deps
  double *lh = nullptr;          // np: adjoint of RHS_T through the next particle temperature update
  double *hpart = nullptr;       // block partials of the exchange's reversal (one per 256 particles)
  // the pull-form transposes of the gather (face fields) and of the exchange's
  // corner gather (lam_Rho): per-particle adjoints, and bins of the positions
  double *lug = nullptr;         // 3 x np: adjoint of the gathered velocity at the stage's positions
  double *gt = nullptr;          // np: adjoint of the gathered gas temperature of an exchange
  int32_t *bkey = nullptr, *bperm = nullptr, *binv = nullptr;  // np each
  int32_t *bstart = nullptr;     // cells + 1
  double *scal_save = nullptr;   // the Update_P0 slots of the run's end, kept over a recomputing sweep
  std::vector<double> I0;        // lamp intensity of each step
  bool swept = false;            // a sweep has run since the recording
  GasAdjoint ga{};
  // steps the gas and sub tapes hold
  int tape_steps() const { return segmented ? std::min(every, nsteps) : nsteps; }
  // the gas tape of step s (4 substeps x 4 fields of n)
  double *gas_of(int s, long n) const {
    if (s < gas_base || s >= gas_base + tape_steps()) throw std::logic_error("adjoint: gas tape step not loaded");
    return gas + (long)(s - gas_base) * 16 * n;
  }
};

// Tile shapes of the two deposit kernels (k_heat_fused, k_vel_planes): x-rows
// per block and planes per z-march (0: the built-in choice), and the online
// tuner that picks them.  Both kernels sum each cell's own particles in bin
// order, so the results do not depend on the shape and the tuner can time
// the candidates inside real steps.
struct TileTune {
  struct Shape { int rows, zc; };
  Shape heat{0, 0}, vel{0, 0};
  bool active = false;
  std::vector<Shape> cand[2];           // 0 heat, 1 velocity
  std::vector<double> best_ms[2];       // fastest time seen per candidate (-1: none yet)
  std::vector<int> runs[2];             // times each candidate was timed
  static constexpr int kRuns = 2;       // each candidate is timed this often before a choice
  size_t next[2] = {0, 0};
  struct Pending { int kind; size_t cand; hipEvent_t a, b; };
  std::vector<Pending> pending;
};

struct Timing {
  bool on = false;
  struct Rec { int cls; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double ms[TPX_K_COUNT] = {};
  int64_t launches[TPX_K_COUNT] = {};
};

}  // namespace tpx

struct tpx_ctx {
  tpx_params p;
  tpx::Geo g;
  double R, P0par, Tp, mp, Vcell;
  double pre[4], post[4];
  double A, gv[3], I0;
  int Iteration;
  double T_cur = 0;
  int num_timestep = 0;
  int num_iteration = 0;
  int64_t total_iterations = 0;    // summed over all Poisson solves
  int iter_guess[4] = {4, 4, 4, 4}; // iterations queued up front, per substep (last solve's count)
  hipStream_t stream = nullptr;
  int device = 0;
  size_t max_lds = 65536;          // dynamic LDS bytes one workgroup may use on this device
  int fft_passes = 3;              // Poisson iteration: 3 passes where every axis allows, else 5
  bool x_packed = true;            // x lengths 64, 128, 256: pass 1 transforms the real rows as half-length complex ones
  tpx::AdjointTape adj;
  // grid fields
  double *ru[5][3] = {};          // buffer sets for RU
  int ru_base = 0, ru_int = 1, ru_new = 2, ru_np1 = 3;
  bool ru_np1_is_base = false;
  double *rho[5] = {};
  int rho_base = 0, rho_int = 1, rho_new = 2, rho_np1 = 3;
  bool rho_np1_is_base = false;
  double *P = nullptr, *rhs = nullptr, *S[3] = {}, *Q = nullptr, *C = nullptr,
         *divergence = nullptr;
  double *dep_tmp = nullptr;      // 3 x n: row partials of the heat deposit
  double *static_scratch = nullptr;  // this device's address of tpx_api.hip's static scratch
  double *coef_stage = nullptr;   // 6 x n: a coefficient solve's C (3 faces), P, RHS and
                                  // residual, allocated on its first call
  // passive scalar (grid.h:79-83): base, int, new and np1 buffers, allocated
  // when it is solved or first uploaded; TimeAdvance_RK4 swaps int and new,
  // TimeAdvance copies np1 into base
  double *ps[4] = {};
  int ps_base = 0, ps_int = 1, ps_new = 2, ps_np1 = 3;
  // passive vector (grid.h:65-70, 99): RV's base, int, new and np1 buffer
  // sets (swapped and copied like the scalar's), its pressure Q, the raw
  // right-hand side of Q's solve, the constant density Rho_forV as a field
  // (the Poisson iteration forms its coefficient from a density) and the
  // type-1 source's filtered field; allocated when the vector is solved or
  // uploaded
  double *rv[4][3] = {};
  int rv_base = 0, rv_int = 1, rv_new = 2, rv_np1 = 3;
  double *Qv = nullptr, *rhs_q = nullptr, *rho_forv_field = nullptr, *rv_filt = nullptr;
  double *track = nullptr;        // tracked particle positions (3 x track_cap)
  int track_cap = 0;
  bool rhs_has_mean = false;      // rhs holds the raw RHS, SC_SUM_RHS its sum
  bool q_zero = false;            // Q is known to be identically zero
  bool div_valid = false;         // divergence holds Compute_Div_U_new's field (the fused
                                  // step forms it inside k_div_rhs and never stores it)
  bool binned_this_step = false;  // particles were re-binned during the current step
  double2 *spec[2] = {};          // half-spectrum ping-pong (nx/2+1)*ny*nz
  double *lam[3] = {};            // modified-wavenumber tables per axis
  double2 *tw[3] = {};            // twiddles per axis (forward)
  double *scal = nullptr;         // device scalars
  double *partials = nullptr;     // block partial results, 3 x n_partials
  bool stats_on = false;          // capture the statistics' last-substep quantities
  tpx::TileTune tile;
  double *st_part = nullptr;      // statistics partials (16 x st_stride) and results
  long st_stride = 0;
  long n_partials = 0;
  double *h_scal = nullptr;       // pinned host mirror of scal
  int32_t *scan_tmp = nullptr;    // scan block sums
  tpx::Particles pt;
  tpx::Timing tm;
  std::string err;
};

namespace tpx {


// ---- launchers (k_grid.hip) ----
void launch_update_rho(tpx_ctx *c, const Coef &k, const double *const ru_int[3],
                       const double *rho_base, const double *rho_np1_in, double *rho_np1,
                       double *rho_new);
void launch_rhs_ru(tpx_ctx *c, const Coef &k, bool with_rho, const double *const ru_int[3],
                   const double *rho_int, const double *const ru_base[3],
                   const double *const ru_np1_in[3], const double *const S[3],
                   double *const ru_np1[3], double *const ru_new[3],
                   const double *rho_base, const double *rho_np1_in, double *rho_np1,
                   double *rho_new);
// mode 0: fused (write rhs + partial sums); 1: write divergence only;
// 2: read divergence, write rhs + partial sums
void launch_div_rhs(tpx_ctx *c, const Coef &k, int mode, const double *const ru_new[3],
                    const double *rho_new, const double *Q, double *divergence, double *rhs);
void launch_project(tpx_ctx *c, const Coef &k, const double *P, double *const ru_np1[3],
                    double *const ru_new[3], bool copy_np1_to_new, bool want_sums);
void launch_shift_mean(tpx_ctx *c, const double *const in[3], double *const out[3],
                       const double a[3]);
void launch_reduce(tpx_ctx *c, const double *f, int op, int slot);  // op 0 sum, 1 max|.|, 2 sum of squares
void launch_finish_sum(tpx_ctx *c, int nparts, int part_offset, int slot);
void launch_update_p0(tpx_ctx *c, const Coef &k);
void launch_fill(tpx_ctx *c, double *f, double v, long n);
void launch_scalar_copy(tpx_ctx *c, int dst, int src);  // scal[dst] = scal[src]
void launch_copy(tpx_ctx *c, double *dst, const double *src, long n);

// ---- launchers (k_fft.hip) ----
int fft_setup(tpx_ctx *c);
// one Poisson iteration; skipped on the device once the solve has stopped
void launch_poisson_iteration(tpx_ctx *c, const double *rho_new);
// the same iteration on caller-given P, raw right-hand side and scalar slots
// sum_slot: the scal slot holding the sum of rhs (its mean is removed in the residual);
// Pout: non-null, P + dp is written there and P is left as it was (an iteration
// skipped on the device then writes nothing)
void launch_poisson_iteration_on(tpx_ctx *c, const double *rho_new, double *P, const double *rhs, double *scal,
                                 const double *res = nullptr, int sum_slot = SC_SUM_RHS, double *Pout = nullptr);
// out = M in, M the Poisson iteration's spectral step (forward transform, the
// division by the modified wavenumbers, inverse, -1/ncell): the iteration's
// dp for the residual `in`.  No loop test runs; scal[SC_DONE] must be 0.
void launch_spectral_apply(tpx_ctx *c, const double *in, double *out, double *scal);
// r = (div_F2C(C grad_C2F P) - RHS) cc for a caller's face coefficient C (cf[0..2])
void launch_residual_coef(tpx_ctx *c, const double *const cf[3], const double *P, const double *rhs, double *res);
// the passive vector (k_vector.hip): Update_RV_WOQ, Compute_RHS_Pois_Q (raw
// right-hand side into rhs_q, its sum into SC_SUM_RHS_Q, with want_vv the sum
// of squares of V into SC_ST_VV), Update_RV_WQ, and the sum of V_Source
void launch_rv_woq(tpx_ctx *c, const Coef &k, const double *const ru_int[3], const double *rho_int);
void launch_rhs_pois_q(tpx_ctx *c, const Coef &k, bool want_vv);
void launch_rv_wq(tpx_ctx *c, const Coef &k);
void launch_rv_source_sum(tpx_ctx *c);
// positions of the particles with flags m*stride, m < count (k_stats.hip)
void launch_track(tpx_ctx *c, int stride, int count, double *out);
// grid::Update_Passive_Scalar (k_grid.hip) and the sum of its source
void launch_passive_scalar(tpx_ctx *c, const Coef &k, const double *const ru_int[3], const double *rho_int);
void launch_passive_scalar_source_sum(tpx_ctx *c);
// resets the device-side iteration count and loop test of a solve
void launch_solve_begin(tpx_ctx *c);
// copies the solve's iteration count and stopped flag to dst[0..1] (device)
void launch_solve_state(tpx_ctx *c, double *dst);

// ---- launchers (k_particle.hip) ----
// count_new: zero start_new and count x_new into it for launch_sort_new
void launch_particle_advance(tpx_ctx *c, const Coef &k, const double *const ru_int[3],
                             const double *rho_int, bool count_new);
void launch_velocity_deposit(tpx_ctx *c);
void launch_sort_new(tpx_ctx *c);
// temperature gather at the stage's particles and heat deposit (writes RHS_T
// by particle and Q)
void launch_heat_exchange(tpx_ctx *c, int stage, const double *rho, int p0_slot);
void launch_concentration(tpx_ctx *c);
void launch_particle_permute(tpx_ctx *c);
void launch_sort_stage(tpx_ctx *c, int stage, bool into_int);
void launch_iota(tpx_ctx *c, int32_t *p, int64_t n);
// Deterministic cell bins of the particles' keys (the particle path's
// binning): start (cells + 1) holds each cell's count on entry and its first
// slot on return, perm the particles in cell order (each cell's by index),
// inv its inverse.  launch_zero_counts clears the cells + 1 counts.
void launch_bin_by_key(tpx_ctx *c, const int32_t *key, int32_t *perm, int32_t *inv, int32_t *start);
void launch_zero_counts(tpx_ctx *c, int32_t *cnt);
void launch_exclusive_scan(tpx_ctx *c, int32_t *data, int64_t n);  // in place, n+1 output

// k_adjoint.hip
void adjoint_tape_substep(tpx_ctx *c, int rk, const double *const ru_int[3], const double *rho_int);
void adjoint_checkpoint(tpx_ctx *c, int slot);
void adjoint_track_permutation(tpx_ctx *c);  // before a re-sort by pt.perm_int during a recording
double adjoint_objective(tpx_ctx *c);  // J = Var(C) of the base positions; leaves dJ/dC in adj.djdc
// The gas half of the reverse sweep as the particle sweep drives it: for each
// step s backwards begin(s), then for each substep rk backwards part_a(s, rk)
// (up to the adjoint of the momentum right-hand side, which it returns: three
// face fields), the particles' stage rk, and part_b(s, rk, lup) with lup the
// particles' dJ/du_g on the faces.  a_drag = -mp/Vcell under two-way
// coupling (the drag deposit's weight in the momentum right-hand side), else 0.
struct GasSweep {
  std::function<void(int)> begin;
  std::function<void(int, int)> part_a;                    // through the Poisson right-hand side (G5, G6)
  std::function<const double *const *(int, int)> part_rk;  // lam_Rho_new complete, RK bookkeeping; returns lam_RHS
  std::function<void(int, int, const double *)> part_b;
  double a_drag = 0.0;
  // heat exchange: where the reversal of the deposit at the end of substep
  // (s, rk) reads and writes, between part_a and part_rk
  struct HeatMid {
    const double *rho;  // Rho_new of the substep
    const double *lq;   // adjoint of its RHS_Part_Temp (null: none)
    double *lrho;       // adjoint of Rho_new, accumulated
  };
  std::function<HeatMid(int, int)> heat_mid;
  HeatMid heat_init{};  // the recorded run's initial exchange (Rho at the start)
  // segment mode: loads the gas tape of checkpoint segment k before the
  // particles' stages of that segment are recomputed (null: the tape is whole)
  std::function<void(int)> segment;
};
// dJ/dtau into adj.dtau and dJ/d(initial particle state) into adj.lam; with
// gas set, the gas sweep runs between the particles' stages
void adjoint_reverse(tpx_ctx *c, const GasSweep *gas);

// k_stats.hip: grid::Statistics.  The captures run in the fourth substep:
// before the particle advance (C and the gathered gas velocity at the stage's
// positions) and after the momentum update (the pre-projection velocity).
void stats_capture_particles(tpx_ctx *c, const double *const ru_int[3], const double *rho_int);
void stats_capture_velocity(tpx_ctx *c, const double *const ru_new[3], const double *rho_new);
void stats_reduce(tpx_ctx *c, const double *rho, const double *const ru[3], double raw[16]);

// k_particle.hip: deposit tile shapes and their tuner
bool tile_shape_valid(const tpx_ctx *c, int kind, int rows, int zc);  // kind 0 heat, 1 velocity
void tile_tune_begin(tpx_ctx *c);
void tile_tune_collect(tpx_ctx *c);  // after a step: read the timings, choose once every candidate ran

// k_adjoint_gas.hip: one substep of the gas sweep in three parts (through
// the Poisson right-hand side; the RK bookkeeping once lam_Rho_new is
// complete; after the particles' stage) and the step boundary.  With the
// heat exchange, part b also reverses the substep's Update_P0.
void adjoint_gas_substep_a(tpx_ctx *c, GasAdjoint &ga, const GasSubstep &s, const Coef &k);
const double *const *adjoint_gas_substep_rk(tpx_ctx *c, GasAdjoint &ga, const GasSubstep &s, const Coef &k);
void adjoint_gas_substep_b(tpx_ctx *c, GasAdjoint &ga, const GasSubstep &s, const Coef &k);
void adjoint_gas_boundary(tpx_ctx *c, GasAdjoint &ga, bool heat);
// Transpose checks of the Poisson adjoint (device arrays of n values; w:
// 8 n + SC_COUNT of work).  op 0: ax = M x, aty = M y (M the spectral step).
// op 1: ax = one iteration from P = x with a zero right-hand side at the
// density rho; aty, atr = that iteration reversed for the adjoint y of its
// output: the adjoint of its input P, and of rho (through C and cc).
void adjoint_poisson_check(tpx_ctx *c, int op, const double *rho, const double *x, const double *y, double *ax,
                           double *aty, double *atr, double *w);

// ---- timing helpers (tpx_api.hip) ----
void timing_begin(tpx_ctx *c, int cls);
void timing_end(tpx_ctx *c, int cls);

inline unsigned grid1d(long n, int block) { return (unsigned)((n + block - 1) / block); }

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X.  Renumber
// them so that each XCD receives one contiguous range of work items, which
// keeps the neighbouring planes a stencil re-reads in that XCD's L2.
__device__ inline unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned nxcd = 8;
  unsigned per = nb / nxcd;
  if (per == 0 || b >= per * nxcd) return b;
  return (b % nxcd) * per + (b / nxcd);
}


// ---- kernel launches: one box, or recorded for a batch of boxes ----------
// Every kernel is a struct F with `static constexpr int kBounds` (its launch
// bound) and `static __device__ void run(...)` (its body).  tpx_launch() starts
// it on a stream; while the launches of a box are being recorded for a batch
// (g_rec set), it stores the launch shape and the arguments instead, and the
// batch later starts one launch per recorded kernel for all of its boxes:
// box = blockIdx.z, arguments from a table in device memory.
// The work-group width (the hidden group-size argument, 12 bytes into the
// implicit kernel arguments).  Kernel bodies are device functions inlined into
// their launch wrappers, where blockDim.x would be read with a run-time test
// for a partial last work group (the compiler drops that test only for reads
// written in the kernel itself); every launch here uses whole work groups.
__device__ __forceinline__ unsigned block_dim_x() {
  return *(const uint16_t *)((const char *)__builtin_amdgcn_implicitarg_ptr() + 12);
}

// A pointer read from an argument table is a generic (flat) address to the
// compiler, where a kernel's own pointer argument is known to be global: its
// accesses become flat loads and stores, which also count against the LDS
// wait counter, so every LDS wait of the body waits on them too.  Every table
// pointer is device (or pinned host) memory: gptr hands it on as a global
// pointer (the empty asm keeps the compiler from folding the two casts away),
// and the accesses compile as for a kernel argument.  An argument struct with
// a to_global() member has its pointers passed through gptr the same way.
template <class V>
__device__ __forceinline__ V *gptr(V *p) {
  __attribute__((address_space(1))) V *g = (__attribute__((address_space(1))) V *)p;
  asm("" : "+s"(g));
  return (V *)g;
}
template <class V, class = void>
struct has_to_global : std::false_type {};
template <class V>
struct has_to_global<V, std::void_t<decltype(std::declval<V &>().to_global())>> : std::true_type {};
template <class V>
__device__ __forceinline__ V *as_global(V *p) {
  return gptr(p);
}
template <class V, std::enable_if_t<has_to_global<V>::value, int> = 0>
__device__ __forceinline__ V as_global(const V &v) {
  V c = v;
  c.to_global();
  return c;
}
template <class V, std::enable_if_t<!has_to_global<V>::value && !std::is_pointer_v<V>, int> = 0>
__device__ __forceinline__ const V &as_global(const V &v) {
  return v;
}

template <class F, class T, size_t... I>
__device__ __forceinline__ void run_args(const T &t, std::index_sequence<I...>) {
  F::run(as_global(std::get<I>(t))...);
}

// Struct arguments reach the body by reference to the kernel argument itself
// (constant memory), so fields the body needs again are reloaded from there
// rather than held in scalar registers.
template <class F, class... A>
__global__ __launch_bounds__(F::kBounds) void k_single(A... a) {
  F::run(a...);
}

template <class F, class... A>
__global__ __launch_bounds__(F::kBounds) void k_batched(const char *__restrict__ tbl, unsigned stride) {
  const auto &t = *reinterpret_cast<const std::tuple<A...> *>(tbl + (size_t)blockIdx.z * stride);
  run_args<F>(t, std::index_sequence_for<A...>{});
}

struct Recorder {
  struct Rec {
    void (*launch)(dim3 grid, dim3 block, size_t shmem, hipStream_t st, const char *tbl, unsigned stride,
                   unsigned nbox);
    dim3 grid, block;
    size_t shmem;
    unsigned size, align;
    size_t off;  // of the argument tuple in bytes
  };
  std::vector<Rec> recs;
  std::vector<char> bytes;
  unsigned nbox = 1;  // boxes of the batch: a launch's blocks are this many times its own grid
  void clear() { recs.clear(); bytes.clear(); }
};
extern thread_local Recorder *g_rec;

template <class F, class... A>
void launch_batched(dim3 grid, dim3 block, size_t shmem, hipStream_t st, const char *tbl, unsigned stride,
                    unsigned nbox) {
  hipLaunchKernelGGL((k_batched<F, A...>), dim3(grid.x, grid.y, nbox), block, shmem, st, tbl, stride);
}

template <class F, class... A>
inline void tpx_launch(dim3 grid, dim3 block, size_t shmem, hipStream_t st, A... a) {
  static_assert((std::is_trivially_copyable<A>::value && ...), "kernel arguments are copied as bytes");
  if (Recorder *r = g_rec) {
    using T = std::tuple<A...>;
    Recorder::Rec rec;
    rec.launch = &launch_batched<F, A...>;
    rec.grid = grid;
    rec.block = block;
    rec.shmem = shmem;
    rec.size = sizeof(T);
    rec.align = alignof(T);
    rec.off = (r->bytes.size() + alignof(T) - 1) / alignof(T) * alignof(T);
    r->bytes.resize(rec.off + sizeof(T));
    const T t(a...);
    memcpy(r->bytes.data() + rec.off, static_cast<const void *>(&t), sizeof(T));
    r->recs.push_back(rec);
    return;
  }
  hipLaunchKernelGGL((k_single<F, A...>), grid, block, shmem, st, a...);
}

}  // namespace tpx

#define TPX_LAUNCH_CHECK()                                  \
  do {                                                      \
    hipError_t e_ = hipGetLastError();                      \
    if (e_ != hipSuccess) throw std::runtime_error(hipGetErrorString(e_)); \
  } while (0)
